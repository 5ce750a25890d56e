/*
 * bh_symplectic.h - the Hamiltonian form of the geodesic on the GPU: batches
 * of phase-space states advanced by the fourth-order Yoshida composition
 * (drop-in for the batch use of src/physics/symplectic_integrator.h:123-133
 * (H), :270-298 (yoshida4Step) and :304-311 (integrateGeodesic)).
 *
 * A state is 8 doubles, (t, r, theta, phi, p_t, p_r, p_theta, p_phi):
 * Boyer-Lindquist coordinates and covariant momenta in geometric units with
 * rS = 2M, H = g^{mu nu} p_mu p_nu / 2 (0 on null geodesics).  Each step
 * applies D(c1) K(w1) D(c2) K(w0) D(c2) K(w1) D(c1), w1 = 1 / (2 - 2^(1/3)),
 * w0 = 1 - 2 w1, c1 = w1 / 2, c2 = (w0 + w1) / 2, where a drift advances q
 * by h dH/dp and a kick advances p by -h dH/dq (analytic derivatives of
 * Carter's separated form), the operations of physics::yoshida4Step
 * (bh_physics.hpp) in the same order, evaluated on the device.
 *
 * Plain C; returns 0 / -1 (message in bh_last_error()).
 */
#ifndef BH_SYMPLECTIC_H
#define BH_SYMPLECTIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* integrateGeodesic for n states at once: nSteps Yoshida steps of
 * dLambda / nSteps each (a negative dLambda integrates backwards), in place.
 * states: device double[n][8]; one state per lane; asynchronous on stream.
 * nSteps <= 0 or n == 0 leaves the states unchanged.  -1 for n < 0, rS <= 0,
 * a non-finite dLambda or a NULL buffer with n > 0. */
int bh_symplectic_integrate(double* states, int64_t n, double dLambda, int32_t nSteps, double a, double rS,
                            void* stream);

/* kerrHamiltonian for n states: H[i] = g^{mu nu} p_mu p_nu / 2 (device
 * buffers, asynchronous on stream). */
int bh_symplectic_hamiltonian(const double* states, int64_t n, double a, double rS, double* H, void* stream);

/* bh_symplectic_integrate over host buffers: copies in, integrates, copies
 * out; synchronous.  For the C++ interface (physics::integrateGeodesicBatch). */
int bh_symplectic_integrate_host(double* states, int64_t n, double dLambda, int32_t nSteps, double a, double rS);

#ifdef __cplusplus
}
#endif

#endif /* BH_SYMPLECTIC_H */
