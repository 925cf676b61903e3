// Process-level runtime shared by the two C-ABIs: device selection, the plan
// cache, factor construction and error reporting.
#pragma once
#include <cstdlib>
#include <memory>

#include "engine.hpp"

namespace gfdfd {

// Select the device once per process: GFDFD_DEVICE if set, else the current
// device.  Several processes (MPI ranks) may share one GPU.  Throws when no
// GPU is present: there is no CPU fallback.
inline void ensure_device() {
    static int selected = -1;
    if (selected >= 0) return;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) throw hip_error("no HIP device available (gfdfd runs on gfx950 only)");
    int dev = 0;
    if (const char *s = std::getenv("GFDFD_DEVICE")) dev = std::atoi(s) % count;
    else GFDFD_HIP(hipGetDevice(&dev));
    GFDFD_HIP(hipSetDevice(dev));
    selected = dev;
}

int report(int code, const char *fmt, ...);
std::shared_ptr<const Plan> cached_host_plan(int nx, int ny, const PlanOptions &opt);
std::shared_ptr<DevPlan> device_plan(std::shared_ptr<const Plan> pl);
std::shared_ptr<DevPlan> cached_grid_plan(int nx, int ny, const PlanOptions &opt);
template <typename T>
std::unique_ptr<Factors<T>> make_factors(std::shared_ptr<DevPlan> dp, int nfreq, const cx<double> *vals, bool copy_vals,
                                         hipStream_t st);

// componentwise backward error of one solve: R = b - op(M) x and
// omega = max_i |r_i| / (|op(M)| |x| + |b|)_i  (UMFPACK's omega1)
double backward_error(const DevPlan &dp, const cx<double> *vals, int sys, const cx<double> *B, const cx<double> *X,
                      cx<double> *R, hipStream_t st);

}  // namespace gfdfd
