// Native SCF loop of the plain form: the per-molecule iteration sequence of
// nanoDFT (_nanoDFT, pyscf_ipu/nanoDFT/nanoDFT.py:141-160, iterating
// nanoDFT_iteration :40-102) with an energy-change stop, issued from C++ so
// that several host threads can drive molecules on their own streams without
// an interpreter lock between launches.  Each iteration: SCF step (Fock,
// DIIS, eigh, density), V_xc and E_xc of the new density, dJK, energy terms,
// then one device-to-host copy of the energies and orbital energies.
#include "common.h"

namespace {

struct RksArgs {
  int dtype, N, nocc, g, use_diis;
  const void* ao;
  const void* weights;
  const void* eri;
  const double *Hcore, *S, *Linv;
  double* D;
  void* Dt;
  void* dJK;
  void* Vxc;
  double *exc, *H, *mo_e, *C, *E;
  void* state;
  void* ws_xc;
  size_t ws_xc_bytes;
  void* ws_jk;
  size_t ws_jk_bytes;
  double e_nuc;
  void* stream;
};

constexpr double EPSILON_B3LYP = 1e-20;  // nanoDFT.py:17
constexpr double HYB_B3LYP = 0.2;        // nanoDFT.py:18

// V_xc, E_xc and dJK of the current density, then the energy terms.
int fock_and_energy(const RksArgs& a) {
  const int NN = a.N * a.N;
  const void* dm = a.D;
  if (a.dtype != PSCF_F64) {
    const int rc = pscf_cast(PSCF_F64, a.D, a.dtype, a.Dt, NN, a.stream);
    if (rc) return rc;
    dm = a.Dt;
  }
  int rc = pscf_xc(a.dtype, a.ao, (int64_t)a.N * a.g, a.g, 1, a.g, a.N, dm, a.weights, EPSILON_B3LYP, a.exc, a.Vxc,
                   a.ws_xc, a.ws_xc_bytes, a.stream);
  if (rc) return rc;
  rc = pscf_jk_s8(a.dtype, a.eri, a.N, dm, HYB_B3LYP, a.dJK, a.ws_jk, a.ws_jk_bytes, a.stream);
  if (rc) return rc;
  return pscf_energy(a.dtype, a.N, a.D, a.Hcore, a.dJK, a.exc, a.e_nuc, a.E, a.stream);
}

// E (5 doubles) and the orbital energies (N) into pinned host memory.
int read_back(const RksArgs& a, double* h) {
  hipStream_t st = (hipStream_t)a.stream;
  PSCF_HIP(hipMemcpyAsync(h, a.E, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
  PSCF_HIP(hipMemcpyAsync(h + 5, a.mo_e, (size_t)a.N * sizeof(double), hipMemcpyDeviceToHost, st));
  PSCF_HIP(hipStreamSynchronize(st));
  return PSCF_OK;
}

}  // namespace

extern "C" int pscf_rks_run(int dtype, int N, int nocc, int g, int max_its, double conv_tol, int use_diis,
                            const void* d_ao, const void* d_weights, const void* d_eri, const double* d_Hcore,
                            const double* d_S, const double* d_Linv, double* d_D, void* d_Dt, void* d_dJK,
                            void* d_Vxc, double* d_exc, double* d_H, double* d_mo_energy, double* d_C,
                            void* d_state, double* d_E, void* d_ws_xc, size_t ws_xc_bytes, void* d_ws_jk,
                            size_t ws_jk_bytes, double e_nuc, double* h_pinned, double* h_log, int* h_iterations,
                            int* h_converged, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_rks_run: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && nocc > 0 && nocc < N && g > 0 && max_its >= 0, "pscf_rks_run: bad sizes");
  PSCF_REQUIRE(d_ao && d_weights && d_eri && d_Hcore && d_S && d_Linv && d_D && d_dJK && d_Vxc && d_exc && d_H &&
                   d_mo_energy && d_C && d_state && d_E && (dtype == PSCF_F64 || d_Dt) && h_pinned && h_log &&
                   h_iterations && h_converged,
               "pscf_rks_run: null pointer");
  const RksArgs a{dtype,   N,     nocc,    g,     use_diis,    d_ao,  d_weights,   d_eri,  d_Hcore,
                  d_S,     d_Linv, d_D,    d_Dt,  d_dJK,       d_Vxc, d_exc,       d_H,    d_mo_energy,
                  d_C,     d_E,   d_state, d_ws_xc, ws_xc_bytes, d_ws_jk, ws_jk_bytes, e_nuc, stream};
  *h_iterations = 0;
  *h_converged = 0;
  int rc = fock_and_energy(a);
  if (rc) return rc;
  if ((rc = read_back(a, h_pinned))) return rc;
  double e_prev = h_pinned[0];
  for (int it = 0; it < max_its; ++it) {
    rc = pscf_scf_step(dtype, it, N, nocc, d_Hcore, d_S, d_Linv, d_dJK, d_Vxc, d_D, d_H, d_mo_energy, d_C, d_state,
                       use_diis, stream);
    if (rc) return rc;
    if ((rc = fock_and_energy(a))) return rc;
    if ((rc = read_back(a, h_pinned))) return rc;
    double* row = h_log + 6 * it;
    for (int k = 0; k < 5; ++k) row[k] = h_pinned[k];
    row[5] = fabs(h_pinned[5 + nocc] - h_pinned[5 + nocc - 1]);
    *h_iterations = it + 1;
    if (fabs(h_pinned[0] - e_prev) < conv_tol) {
      *h_converged = 1;
      break;
    }
    e_prev = h_pinned[0];
  }
  return PSCF_OK;
}
