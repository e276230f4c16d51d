// mps_motion.hip -- prescribed wall motion (a sinusoidal piston).
//
// The reference declares a MOTION class whose doMotion is empty
// (MPS_CAL_/MOTION.cpp:11-16); this fills that role for the wave tank: the
// flagged particles (indexed by particle id, so the cell sort does not
// matter) are placed at x_ref + A sin(w t) along one axis with velocity
// A w cos(w t), t = the end time of the step about to run (the dt kernel has
// already advanced DevScalars::time to it).
#include "mps_device.hpp"
#include "mps_kernels.hpp"

namespace mps {

// x_ref[id] = x_axis - A sin(w time): the law then continues from the current
// position; the velocity takes its value at the current time.
template <typename Real>
__global__ void __launch_bounds__(256) k_motion_ref(int n, StateView<Real> s, const unsigned char* __restrict__ mov,
                                                    double* __restrict__ xref, int axis, double amp, double omega,
                                                    const DevScalars* __restrict__ sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = s.id[i];
    if (!mov[o]) return;
    const Real* xa = axis == 0 ? s.x : (axis == 1 ? s.y : s.z);
    xref[o] = (double)xa[i] - amp * sin(omega * sc->time);
    const Real va = (Real)(amp * omega * cos(omega * sc->time));
    s.u[i] = axis == 0 ? va : (Real)0;
    s.v[i] = axis == 1 ? va : (Real)0;
    s.w[i] = axis == 2 ? va : (Real)0;
}

template <typename Real>
__global__ void __launch_bounds__(256) k_motion(int n, StateView<Real> s, const unsigned char* __restrict__ mov,
                                                const double* __restrict__ xref, int axis, double amp,
                                                double omega, const DevScalars* __restrict__ sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = s.id[i];
    if (!mov[o]) return;
    const double t = sc->time;
    Real* xa = axis == 0 ? s.x : (axis == 1 ? s.y : s.z);
    xa[i] = (Real)(xref[o] + amp * sin(omega * t));
    const Real va = (Real)(amp * omega * cos(omega * t));
    s.u[i] = axis == 0 ? va : (Real)0;
    s.v[i] = axis == 1 ? va : (Real)0;
    s.w[i] = axis == 2 ? va : (Real)0;
}

template <typename Real>
void launch_motion_ref(int n, StateView<Real> s, const unsigned char* mov, double* xref, int axis, double amp,
                       double omega, const DevScalars* sc, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_motion_ref<Real>, dim3((n + 255) / 256), dim3(256), 0, st, n, s, mov, xref, axis, amp,
                       omega, sc);
}

template <typename Real>
void launch_motion(int n, StateView<Real> s, const unsigned char* mov, const double* xref, int axis, double amp,
                   double omega, const DevScalars* sc, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_motion<Real>, dim3((n + 255) / 256), dim3(256), 0, st, n, s, mov, xref, axis, amp, omega,
                       sc);
}

template void launch_motion_ref<float>(int, StateView<float>, const unsigned char*, double*, int, double, double,
                                       const DevScalars*, hipStream_t);
template void launch_motion_ref<double>(int, StateView<double>, const unsigned char*, double*, int, double,
                                        double, const DevScalars*, hipStream_t);
template void launch_motion<float>(int, StateView<float>, const unsigned char*, const double*, int, double,
                                   double, const DevScalars*, hipStream_t);
template void launch_motion<double>(int, StateView<double>, const unsigned char*, const double*, int, double,
                                    double, const DevScalars*, hipStream_t);

}  // namespace mps
