/*
 * C-ABI of libpyscf_mi355x.so: MI355X (gfx950) kernels for the restricted
 * Kohn-Sham B3LYP Fock build and SCF loop of pyscf-ipu's nanoDFT.
 *
 * Conventions
 *   - Every pointer argument named d_* is device memory owned by the caller;
 *     the library never frees caller memory.  h_* pointers are host memory.
 *   - `stream` is a hipStream_t (NULL = default stream); all calls are
 *     asynchronous on it unless stated otherwise.
 *   - Return value: PSCF_OK (0) or a negative PSCF_E* code; the message of the
 *     last error on the calling thread is returned by pscf_last_error().
 *   - dtype selects the arithmetic of the floating-point operands:
 *     PSCF_F64 (double) or PSCF_F32 (float).
 *   - Molecules use libcint's atm/bas/env layout (atm: natm x 6 int32,
 *     bas: nbas x 8 int32, env: double), s and p shells only (l <= 1).
 *   - Workspaces (d_ws, ws_bytes) are caller-provided scratch; the matching
 *     *_workspace_bytes() query returns the size needed.
 */
#ifndef PYSCF_MI355X_H
#define PYSCF_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSCF_OK 0
#define PSCF_EINVAL -1
#define PSCF_EHIP -2
#define PSCF_EUNSUPPORTED -3

#define PSCF_F64 0
#define PSCF_F32 1

const char* pscf_last_error(void);
int pscf_abi_version(void);

/* ---------------------------------------------------------------- ERIs ----
 * pscf_eri_s8: all unique spherical ERIs (ij|kl), p>=q, r>=s, pq>=rs, packed as
 * d_out[pq*(pq+1)/2 + rs] (PySCF aosym="s8" layout, PySCF 2.2.1
 * `mol.intor("int2e_sph", aosym="s8")` as called at
 * pyscf_ipu/nanoDFT/nanoDFT.py:285).  The caller passes the shell pairs
 * (i, j) sorted by pair class (ss, then ps with the p shell first, then pp) in
 * d_pairs (npair x 2 int32) and the three class counts; every s8 element of the
 * basis is written exactly once.  d_out has nao(nao+1)/2 * (nao(nao+1)/2 + 1)/2
 * elements of dtype.
 */
int pscf_eri_s8(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env,
                const int* d_pairs, const int* h_class_counts, int nao, void* d_out, void* stream);

/* pscf_eri_quartets: contracted shell quartets of an explicit list, the work of
 * pyscf_ipu/electron_repulsion/direct.py:513 `electron_repulsion_integrals`
 * (int2e_sph.cpp:6750 per quartet).  Quartet q = d_quartets[4q..4q+3] writes
 * di*dj*dk*dl values at d_out + d_offsets[q], element (a,b,c,d) at
 * a + di*(b + dj*(c + dk*d)) (libcint's column-major block order). */
int pscf_eri_quartets(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env,
                      const int* d_quartets, const int64_t* d_offsets, int nq, void* d_out, void* stream);

/* pscf_int1e: overlap S, kinetic T and nuclear attraction V (all nuclei) in
 * FP64, nao x nao row-major each (the role of mol.intor_symmetric('int1e_ovlp'
 * /'int1e_kin'/'int1e_nuc'), pyscf_ipu/nanoDFT/nanoDFT.py:236-238). */
int pscf_int1e(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, int nao,
               double* d_S, double* d_T, double* d_V, void* stream);

/* ----------------------------------------------------------------- J/K ----
 * pscf_jk_sparse: drop-in for sparse_symmetric_einsum
 * (pyscf_ipu/nanoDFT/sparse_symmetric_ERI.py:86-137 with the index table of
 * compute_indices.cpp:54-95): d_out (N x N) = sum over the 8 J patterns minus
 * hyb/2 times the 8 K patterns of gather(dm) * eri -> scatter-add, for nnz
 * ERIs already divided by their permutation multiplicity, with int16 AO
 * indices (i, j, k, l) per ERI (d_idx: nnz x 4).  dm need not be symmetric.
 */
size_t pscf_jk_sparse_workspace_bytes(int dtype, int N, int64_t nnz);
int pscf_jk_sparse(int dtype, const void* d_eri, const int16_t* d_idx, int64_t nnz, const void* d_dm, int N,
                   double hyb, void* d_out, void* d_ws, size_t ws_bytes, void* stream);

/* pscf_jk_s8: the same dJK = J - hyb/2 K from the packed s8 array of
 * pscf_eri_s8 (raw ERIs, not divided by multiplicity), indices regenerated
 * from each element's position; dm must be symmetric.  Equal to
 * pscf_jk_sparse on the nonzero subset. */
size_t pscf_jk_s8_workspace_bytes(int dtype, int N);
int pscf_jk_s8(int dtype, const void* d_eri_s8, int N, const void* d_dm, double hyb, void* d_out, void* d_ws,
               size_t ws_bytes, void* stream);

/* pscf_jk_s8_noalloc: pscf_jk_s8 without a workspace argument and without any
 * allocation, for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The partials live in a library-owned device
 * array of pscf_jk_s8_noalloc_max_bytes() bytes (32 MiB) per GPU, held until
 * the process ends; a call whose pscf_jk_s8_workspace_bytes(dtype, N) exceeds
 * it returns PSCF_EINVAL.  Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or events).  Results equal pscf_jk_s8's. */
size_t pscf_jk_s8_noalloc_max_bytes(void);
int pscf_jk_s8_noalloc(int dtype, const void* d_eri_s8, int N, const void* d_dm, double hyb, void* d_out,
                       void* stream);

/* pscf_jk_blocks: vj and vk of ipu_einsum(__out3, dm, mol)
 * (pyscf_ipu/electron_repulsion/direct.py:489-504, vertices
 * poplar_direct_s1*, int2e_sph.cpp:7598-8176): vj[k][l] = sum (ij|kl) dm[j][i],
 * vk[i][l] = sum (ij|kl) dm[j][k] over the ERI blocks of the shell-quartet list
 * (electron_repulsion_integrals' output concatenated: nclass classes of
 * uniform block size h_class_size[c], class c's elements starting at
 * h_class_elem0[c], its blocks at h_class_blk0[c]; h_class_elem0[nclass] =
 * total elements).  d_blocks: int32 [nblk][4] per block = (di | dj << 4 |
 * dk << 8 | dl << 12 | permutation mask << 16, i0 | j0 << 16, k0 | l0 << 16,
 * 0); block elements in libcint's column-major (di, dj, dk, dl) order; mask bit
 * p applies the p-th index image of the reference's do_list
 * (direct.py:1029-1110).  dm need not be symmetric.  LDS holds dm, vj and vk
 * (3 N^2 elements of dtype, at most 160 KiB). */
size_t pscf_jk_blocks_workspace_bytes(int dtype, int N, int64_t nelem);
int pscf_jk_blocks(int dtype, const void* d_eri, const int* d_blocks, int nclass, const int64_t* h_class_elem0,
                   const int* h_class_blk0, const int* h_class_size, const void* d_dm, int N, void* d_vj, void* d_vk,
                   void* d_ws, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------- XC ----
 * pscf_ao_eval: AO values and gradients on grid points, the role of
 * mol.eval_gto('GTOval_sph_deriv1', coords, 4) (nanoDFT.py:219-220).
 * d_coords: g x 3 double (Bohr).  d_ao: 4 x nao x g of dtype (component,
 * AO, point; point index fastest). */
int pscf_ao_eval(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, int nao,
                 const double* d_coords, int g, void* d_ao, void* stream);

/* pscf_becke_weights: per-point Becke fuzzy-cell weights
 * w = vol * P_owner / sum_i P_i with Becke's three-times-smoothed step and the
 * Treutler size adjustment table d_adjust (natm x natm, a_ij = -a_ji), the
 * partition step of PySCF's Grids.build (pyscf_ipu/pyscf_utils/build_grid.py:
 * 34-100, gen_grid_partition).  d_coords: g x 3, d_owner: g atom indices,
 * d_vol: g un-partitioned quadrature weights, d_atom_xyz: natm x 3 (Bohr). */
int pscf_becke_weights(const double* d_coords, const int* d_owner, const double* d_vol, int g,
                       const double* d_atom_xyz, const double* d_adjust, int natm, double* d_w, void* stream);

/* pscf_b3lyp: drop-in for b3lyp(rho, EPSILON_B3LYP)
 * (pyscf_ipu/exchange_correlation/b3lyp.py:13-39): d_rho is 4 x g
 * (rho, grad rho / 2), outputs per point e_xc (energy per particle),
 * v_rho = dE/drho and v_norms = dE/d|grad rho|^2. */
int pscf_b3lyp(int dtype, const void* d_rho, int g, double eps, void* d_exc, void* d_vrho, void* d_vnorm,
               void* stream);

/* pscf_b3lyp_host: the same pointwise B3LYP code evaluated on the host in
 * FP64 (h_* host arrays, same layout); no GPU needed. */
int pscf_b3lyp_host(const double* h_rho, int g, double eps, double* h_exc, double* h_vrho, double* h_vnorm);

/* pscf_xc: drop-in for exchange_correlation(dm, grid_AO, grid_weights)
 * (pyscf_ipu/nanoDFT/nanoDFT.py:105-123).  d_ao holds AO values/gradients
 * with element (c, mu, p) at d_ao[c*ao_sc + mu*ao_smu + p*ao_sp] (so both the
 * reference's (4, g, N) layout and pscf_ao_eval's (4, N, g) layout work).
 * Outputs: *d_exc (one double, E_xc) and d_vxc (N x N, dtype, V + V^T). */
size_t pscf_xc_workspace_bytes(int dtype, int N, int g);
int pscf_xc(int dtype, const void* d_ao, int64_t ao_sc, int64_t ao_smu, int64_t ao_sp, int g, int N,
            const void* d_dm, const void* d_weights, double eps, double* d_exc, void* d_vxc, void* d_ws,
            size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------ SCF ----
 * pscf_scf_step: one nanoDFT_iteration up to the new density
 * (pyscf_ipu/nanoDFT/nanoDFT.py:64-72 with DIIS :338-364 and linalg_eigh
 * :370-393), FP64, one workgroup:
 *   H = H_core + dJK + V_xc; DIIS(9) extrapolation (errvec = S D H - (S D H)^T,
 *   c = pinv(B)[0]); eigh(L^-1 H L^-T) by warm-started parallel Jacobi;
 *   C = L^-T U; D = 2 C_occ C_occ^T.
 * d_state: pscf_scf_state_bytes(N) bytes, zero-filled before iteration 0 by
 * the caller, carried between calls (DIIS history, previous eigenvectors).
 * dJK / V_xc are of dtype; everything else FP64.  Outputs d_D (N x N),
 * d_mo_energy (N), d_C (N x N); the H actually diagonalised (after DIIS) is
 * written to d_H (N x N). */
size_t pscf_scf_state_bytes(int N);
int pscf_scf_step(int dtype, int iteration, int N, int nocc, const double* d_Hcore, const double* d_S,
                  const double* d_Linv, const void* d_dJK, const void* d_Vxc, double* d_D, double* d_H,
                  double* d_mo_energy, double* d_C, void* d_state, int use_diis, void* stream);

/* pscf_energy: E = sum D*H_core + 1/2 sum D*dJK + E_xc + E_nuc (nanoDFT.py:20-38)
 * in FP64 with compensated summation; d_out receives
 * [E, E_core, E_JK, E_xc, E_nuc] (5 doubles). */
int pscf_energy(int dtype, int N, const double* d_D, const double* d_Hcore, const void* d_dJK,
                const double* d_exc, double e_nuc, double* d_out, void* stream);

/* pscf_rks_run: the plain form's whole SCF loop for one molecule (_nanoDFT,
 * nanoDFT.py:141-160 over nanoDFT_iteration :40-102), issued from C++:
 * V_xc/E_xc, dJK and energy of the initial density d_D, then up to max_its
 * times pscf_scf_step, pscf_xc, pscf_jk_s8 and pscf_energy on `stream`,
 * stopping when |E_i - E_(i-1)| < conv_tol (a negative conv_tol runs all
 * iterations, like the reference).  d_Dt is the density in dtype (FP32 runs;
 * ignored in FP64), d_ao the (4, N, g) AO block, d_state zero-filled
 * (pscf_scf_state_bytes).  h_pinned: 5 + N doubles of pinned host memory for
 * the per-iteration read-back; h_log receives max_its rows of
 * [E, E_core, E_JK, E_xc, E_nuc, |e_LUMO - e_HOMO|]; *h_iterations the rows
 * written, *h_converged 1 if the stop criterion was met. */
int pscf_rks_run(int dtype, int N, int nocc, int g, int max_its, double conv_tol, int use_diis, const void* d_ao,
                 const void* d_weights, const void* d_eri, const double* d_Hcore, const double* d_S,
                 const double* d_Linv, double* d_D, void* d_Dt, void* d_dJK, void* d_Vxc, double* d_exc, double* d_H,
                 double* d_mo_energy, double* d_C, void* d_state, double* d_E, void* d_ws_xc, size_t ws_xc_bytes,
                 void* d_ws_jk, size_t ws_jk_bytes, double e_nuc, double* h_pinned, double* h_log, int* h_iterations,
                 int* h_converged, void* stream);

/* ------------------------------------------------------ nuclear gradient ----
 * grad(mol, coords, weight, mo_coeff, mo_energy) (pyscf_ipu/nanoDFT/nanoDFT.py:
 * 451-477, grad_elec :403-433) without forming its integral tensors: each
 * entry contracts one class of derivative integrals with the densities as it
 * generates them and ADDS into d_grad (natm x 3 doubles, Ha/Bohr, zeroed by
 * the caller).  d_dm = dm0 = 2 C_occ C_occ^T, d_dme = dme0 = 2 C_occ e_occ
 * C_occ^T (N x N, FP64).  Atomic adds of contributions rounded to 2^-40, so
 * the sums are exact and repeat bit for bit.  The nuclear repulsion term
 * (grad_nuc :435-449) is the caller's.  FP64 only. */
/* h1aos . dm0 and -2 s1 . dme0 (int1e_ipkin, int1e_ipnuc, int1e_iprinv,
 * int1e_ipovlp; nanoDFT.py:456-470): over ordered shell pairs. */
int pscf_grad_int1e(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, int nao,
                    const double* d_dm, const double* d_dme, double* d_grad, void* stream);
/* vj - hyb/2 vk of int2e_ip1 contracted with dm0 (nanoDFT.py:416-418, :428):
 * the unique shell quartets of pscf_eri_s8 (same class-sorted pair list). */
int pscf_grad_eri(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, const int* d_pairs,
                  const int* h_class_counts, int nao, const double* d_dm, double hyb, double* d_grad, void* stream);
/* The B3LYP part (vmat and the deriv-2 AO term, nanoDFT.py:407-415) on the
 * SCF's grid (coords g x 3 Bohr, weights g), AO values and derivatives
 * evaluated in the kernel; workspace pscf_grad_xc_workspace_bytes. */
size_t pscf_grad_xc_workspace_bytes(int nao, int nbas, int g);
int pscf_grad_xc(const double* d_coords, const double* d_weights, int g, const int* d_atm, int natm, const int* d_bas,
                 int nbas, const double* d_env, int nao, const double* d_dm, double eps, void* d_ws, size_t ws_bytes,
                 double* d_grad, void* stream);

/* --------------------------------------------------------- batched form ----
 * Many molecules of uneven size per launch.  d_mols is an int64 [nmol][16]
 * descriptor array (natm, nbas, nao, nocc, atm0, bas0, ns, np, shl0, ngrid,
 * grid0, eri0, mat0, vec0, ao0, state0) giving each molecule's offsets into
 * the concatenated buffers: global atm/bas/env arrays (bas atom indices and
 * env pointers absolute), d_shell_ao (each shell's first AO within its
 * molecule), d_shl_sorted (per molecule: s shells then p shells), packed s8
 * ERIs at eri0, N x N matrices at mat0, length-N vectors at vec0, AO blocks
 * (4 x N x ngrid) at ao0 and SCF state at state0.  d_ctl is double [nmol][8]
 * (iteration, active, converged (-1 = failed set-up), E_prev, E_last, gap,
 * iterations used, spare); kernels skip molecules whose active flag is 0.
 * d_order lists molecules largest first (load balance of one workgroup per
 * molecule).  The per-molecule arithmetic is that of the single-molecule
 * entry points above; the reference runs one molecule per process
 * (pyscf_ipu/dft.py:240-245, generate.sh:11-14) through the same
 * nanoDFT_iteration (nanoDFT.py:40-102).
 */
/* ERIs of all molecules: h_work[c] (device, [h_nwork[c]][2] int32 of
 * (molecule, first quartet)) for the classes (ss|ss), (ps|ss), (ps|ps),
 * (pp|ss), (pp|ps), (pp|pp); mol.intor("int2e_sph", aosym="s8") per molecule
 * (nanoDFT.py:285). */
int pscf_batch_eri_s8(int dtype, const void* d_mols, int nmol, const int* const* h_work, const int* h_nwork,
                      const int* d_atm, const int* d_bas, const double* d_env, const int* d_shell_ao,
                      const int* d_shl_sorted, void* d_out, void* stream);
/* pscf_batch_eri_s8 with every element divided by its multiplicity
 * 2^([p==q] + [r==s] + [pq==rs]) as it is written, the form
 * pscf_batch_eri_scale produces (sparse_symmetric_ERI.py:68-79, 180-182),
 * without the second pass over the ERIs. */
int pscf_batch_eri_s8_div(int dtype, const void* d_mols, int nmol, const int* const* h_work, const int* h_nwork,
                          const int* d_atm, const int* d_bas, const double* d_env, const int* d_shell_ao,
                          const int* d_shl_sorted, void* d_out, void* stream);
/* Work list (owner, first index) pairs [total][2] covering range(count_i) of
 * each owner i in steps of `size`, from d_first = the exclusive prefix sums
 * of ceil(count_i / size) (n + 1 entries, d_first[n] = total): the chunk
 * lists of pscf_batch_grid and pscf_batch_eri_s8, built on the device. */
int pscf_batch_chunks(const int64_t* d_first, int n, int size, int* d_work, int64_t total, void* stream);
/* Nuclear repulsion energy of every molecule (mol.energy_nuc(), nanoDFT.py). */
int pscf_batch_enuc(const void* d_mols, int nmol, const int* d_atm, const double* d_env, double* d_enuc, void* stream);
/* S and H_core = T + V of all molecules (nanoDFT.py:236-238). */
int pscf_batch_int1e(const void* d_mols, int nmol, const int* d_atm, const int* d_bas, const double* d_env,
                     const int* d_shell_ao, double* d_S, double* d_Hcore, void* stream);
/* L^-1 = inv(cholesky(S)) per molecule (nanoDFT.py:239); a non-positive-definite
 * S marks the molecule failed and inactive (the skip of pyscf_ipu/dft.py:457-464). */
int pscf_batch_chol(const void* d_mols, int nmol, int max_nao, const double* d_S, double* d_Linv, double* d_ctl,
                    void* stream);
/* Atomic-density initial guess per molecule (role of init_guess_by_minao,
 * nanoDFT.py:233); d_Dt receives the same density in dtype. */
int pscf_batch_guess(int dtype, const void* d_mols, int nmol, const int* d_atm, const int* d_bas,
                     const int* d_shell_ao, const double* d_S, double* d_D, void* d_Dt, void* stream);
/* Grid points of all atoms from per-element templates (Grids.build atom grids,
 * nanoDFT.py:215-217): d_work [nwork][2] = (atom, first template point). */
int pscf_batch_grid(const int* d_work, int nwork, const int* d_atom_elem, const int64_t* d_atom_pt0,
                    const int* d_atom_mol, const int* d_tmpl_off, const int* d_tmpl_n, const double* d_tmpl,
                    const int* d_atm, const double* d_env, double* d_coords, double* d_vol, int* d_owner,
                    int* d_ptmol, void* stream);
/* Becke weights of all points (build_grid.py:34-100); d_w64 optional FP64 copy. */
int pscf_batch_becke(int dtype, const void* d_mols, const int* d_ptmol, const int* d_owner, const double* d_coords,
                     const double* d_vol, int64_t G, const int* d_atm, const double* d_env, void* d_w, double* d_w64,
                     void* stream);
/* AO values and gradients of all points (eval_gto, nanoDFT.py:219-220). */
int pscf_batch_ao(int dtype, const void* d_mols, const int* d_ptmol, const double* d_coords, int64_t G,
                  const int* d_atm, const int* d_bas, const double* d_env, void* d_ao, void* stream);
/* E_xc (d_exc[m]) and V_xc of the active molecules d_order[0..nmol)
 * (nanoDFT.py:105-123), grid tiles of `tile` points (32, or 64 in FP32);
 * max_nao bounds the nao of these molecules (it sizes the LDS and the tile
 * shapes; a larger molecule gets NaN).  Each molecule's grid is cut into
 * chunks of PSCF_XC_CHUNK_PTS points whose V and E_xc partials are summed in
 * chunk order, so the result does not depend on `split` (workgroups sharing a
 * molecule's chunks).  Chunk partials go to d_part_v (part_v_elems >=
 * nchunk_max * nmol * max_nao^2 elements of dtype) and d_part_e (part_e_elems
 * >= nchunk_max * nmol doubles), nchunk_max = ceil(largest grid /
 * PSCF_XC_CHUNK_PTS); a molecule whose chunks do not fit gets NaN. */
#define PSCF_XC_CHUNK_PTS 512
int pscf_batch_xc(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao, const double* d_ctl,
                  const void* d_ao, const void* d_dm, const void* d_w, double eps, void* d_vxc, double* d_exc,
                  int split, int tile, void* d_part_v, double* d_part_e, int64_t part_v_elems, int64_t part_e_elems,
                  void* stream);
/* pscf_batch_xc with the AO values and gradients evaluated inside the kernel
 * (eval_gto, nanoDFT.py:219-220, fused into exchange_correlation,
 * nanoDFT.py:105-123) instead of read from a stored AO block: each
 * workgroup tabulates its molecule's s and p shells (d_atm/d_bas/d_env as
 * given to pscf_batch_ao) in LDS and evaluates the tile's points from
 * d_coords [G][3].  nsh_max >= the shells of any molecule of the call (at most
 * 80), kp >= the primitives of any shell (at most 16); a molecule with more
 * shells gets NaN.  Same results as pscf_batch_ao + pscf_batch_xc up to
 * rounding, without the 4 x nao x ngrid AO block in HBM. */
int pscf_batch_xc_direct(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                         const double* d_ctl, const double* d_coords, const int* d_atm, const int* d_bas,
                         const double* d_env, int nsh_max, int kp, const void* d_dm, const void* d_w, double eps,
                         void* d_vxc, double* d_exc, int split, int tile, void* d_part_v, double* d_part_e,
                         int64_t part_v_elems, int64_t part_e_elems, void* stream);
/* Divides every molecule's packed s8 ERIs (pscf_batch_eri_s8) in place by
 * their multiplicity 2^([p==q] + [r==s] + [pq==rs]), as the reference does
 * before its contraction (sparse_symmetric_ERI.py:68-79, 180-182); max_nao
 * bounds the molecules' nao. */
int pscf_batch_eri_scale(int dtype, const void* d_mols, int nmol, int max_nao, void* d_eri, void* stream);
/* dJK of the active molecules d_order[0..nmol) from their packed s8 ERIs
 * divided by their multiplicity (pscf_batch_eri_scale)
 * (sparse_symmetric_ERI.py:86-137); max_nao bounds their nao (it sizes the
 * LDS; a larger molecule gets NaN).  accumulate = 0: d_out (dtype) =
 * dJK(d_dm); accumulate = 1: d_out (FP64) += dJK(d_dm), the incremental build
 * from density changes.  split workgroups of `threads` (256 or 512) threads
 * share each molecule's ERI rows (atomic accumulation).  Each workgroup holds
 * dm (dtype) and an FP64 accumulator in LDS: max_nao^2 * (sizeof(dtype) + 8)
 * bytes, at most 160 KiB (max_nao <= 116 in FP32, 101 in FP64). */
int pscf_batch_jk_s8(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao, const double* d_ctl,
                     const void* d_eri, const void* d_dm, double hyb, void* d_out, int accumulate, int split,
                     int threads, void* stream);
/* One SCF step of all active molecules (nanoDFT.py:64-72, 338-393); iteration
 * index per molecule from d_ctl; dJK in FP64 (written or accumulated by
 * pscf_batch_jk_s8), V_xc in dtype; writes D (FP64), D in dtype (d_Dt) and,
 * if d_dDt is not NULL, the change of D in dtype (for the incremental J/K). */
int pscf_batch_scf_step(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao, double* d_ctl,
                        const double* d_Hcore, const double* d_S, const double* d_Linv, const double* d_dJK,
                        const void* d_Vxc, double* d_D, void* d_Dt, void* d_dDt, double* d_H, double* d_mo_energy,
                        double* d_C, double* d_state, int use_diis, void* stream);
/* Energies (d_E [nmol][5], nanoDFT.py:20-38) and convergence control: init
 * records E_prev; otherwise a molecule converges when |E - E_prev| < tol and
 * stops at max_its iterations.  dJK in FP64.  d_hist (optional, [nmol][5]
 * doubles): ring of the last five iteration energies, slot = iteration mod 5
 * (the initial energy is iteration 0), for the convergence spread the QM1B
 * rows record (std of the last five energies, pyscf_ipu/dft.py:631). */
int pscf_batch_energy(int dtype, const void* d_mols, int nmol, const double* d_e_nuc, const double* d_D,
                      const double* d_Hcore, const double* d_dJK, const double* d_exc, const double* d_mo_energy,
                      double* d_ctl, double* d_E, double* d_hist, int init, double tol, int max_its, void* stream);

/* ---- large molecules and the precision study ---------------------------------
 * For 6-31G molecules (nao 130-200) whose matrices exceed the LDS-resident
 * kernels.  Same quantities as the calls above, other data paths. */

/* The full s1 ERI tensor T[p][q][r][s] = (pq|rs) (N^4 elements of dtype_out)
 * from the packed s8 array (PySCF aosym="s8", nanoDFT.py:285). */
int pscf_eri_s8_to_s1(int dtype_in, const void* d_s8, int dtype_out, void* d_s1, int N, void* stream);
/* dJK = J - hyb/2 K (FP64, N x N) from the s1 tensor and an FP64 density
 * (the einsum definitions at nanoDFT.py:130-132), one pass over the tensor. */
int pscf_jk_s1(int dtype, const void* d_s1, int N, const double* d_dm, double hyb, double* d_out, void* stream);
/* dJK = J - hyb/2 K (FP64, N x N) from the packed s8 ERIs (raw values, as
 * pscf_eri_s8 writes them; sparse_symmetric_ERI.py:86-137) and an FP64
 * density, for nao beyond the LDS-resident pscf_jk_s8 (which holds dm and an
 * N x N accumulator in LDS: nao <= 116 in FP32, 101 in FP64): only rows of dm
 * and of the accumulator sit in LDS, the accumulator itself in the workspace
 * (pscf_jk_s8_large_workspace_bytes(N): N^2 + N(N+1)/2 doubles).  Reads the
 * s8 array twice instead of streaming the N^4 s1 tensor (pscf_jk_s1).
 * Results repeat bit for bit (contributions rounded to multiples of 2^-40
 * before their atomic sums).  N <= 1280. */
size_t pscf_jk_s8_large_workspace_bytes(int N);
int pscf_jk_s8_large(int dtype, const void* d_eri_s8, int N, const double* d_dm, double hyb, double* d_out,
                     void* d_ws, size_t ws_bytes, void* stream);
/* General GEMM C (FP64) = op(A) op(B) on MFMA: inputs stored as in_type
 * (0 BF16, 1 FP16, 2 FP32, 3 FP64), accumulation FP64 (acc64) or FP32;
 * row-major with leading dimensions; atomic = add into C (K split over
 * ksplit workgroups). */
int pscf_gemm(int in_type, int acc64, int trans_a, int trans_b, int atomic, int M, int N, int K, const void* d_A,
              int64_t lda, const void* d_B, int64_t ldb, double* d_C, int64_t ldc, int ksplit, void* stream);
/* FP64 -> out_type (0 BF16, 1 FP16, 2 FP32, 3 FP64), round to nearest. */
int pscf_cast_to(int out_type, const double* d_x, void* d_y, int64_t n, void* stream);
/* Pointwise stages of exchange_correlation (nanoDFT.py:105-123) around the
 * AO-grid GEMMs X = D phi0 and V = phi0 Y^T, AO values phi [4][N][g] stored as
 * ao_type:  stage 0: rho_c = sum_nu X phi_c, B3LYP (b3lyp.py:13-39), E_xc
 * block partials (pscf_xc_large_parts(g) of them), coefficients coef [4][g];
 * stage 1: Y = sum_c coef_c phi_c (ao_type).  finish: V_xc = Vh + Vh^T and
 * E_xc from the partials. */
int pscf_xc_large_parts(int g);
int pscf_xc_large_stage(int stage, int ao_type, int N, int g, const double* d_X, const void* d_phi, const double* d_w,
                        double eps, double* d_coef, double* d_e_part, void* d_Y, void* stream);
int pscf_xc_large_finish(int N, const double* d_Vh, const double* d_e_part, int nparts, double* d_vxc, double* d_exc,
                         void* stream);
/* pscf_scf_step with the working matrices in a caller workspace of
 * pscf_scf_step_ws_bytes(N) bytes (N <= 512). */
size_t pscf_scf_step_ws_bytes(int N);
int pscf_scf_step_ws(int dtype, int iteration, int N, int nocc, const double* d_Hcore, const double* d_S,
                     const double* d_Linv, const void* d_dJK, const void* d_Vxc, double* d_D, double* d_H,
                     double* d_mo_energy, double* d_C, void* d_state, int use_diis, void* d_ws, size_t ws_bytes,
                     void* stream);

/* ---- learned initial guess --------------------------------------------------
 * No counterpart in the reference, whose SCF always starts from PySCF's MINAO
 * guess (nanoDFT.py:233); these replace that call when a trained network is
 * given, and MINAO stays the fallback.  Per ordered atom pair an MLP
 * 43 -> 256 -> 256 -> 256 -> 25 (SiLU) predicts the STO-3G density block. */

/* Symmetrise and purify predicted densities: D <- 2 L^-T V_occ V_occ^T L^-1
 * with V the nocc eigenvectors of largest occupation of L^T D L. */
int pscf_batch_purify(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao, const double* d_ctl,
                      const double* d_S, const double* d_Linv, double* d_D, void* d_Dt, void* stream);
/* Pair features: d_pairs [P][2] global atom indices; d_X16 [P][64] BF16 (43
 * used: one-hot H/C/N/O of a and b, 32 Gaussian RBFs of |R_b - R_a| on 0..5
 * Angstrom with sigma 0.15, unit direction a->b); optional d_X32 FP32 copy;
 * d_valid[p] = distance < 5 Angstrom. */
int pscf_pair_features(int64_t P, const int* d_pairs, const int* d_atm, const double* d_env, void* d_X16,
                       float* d_X32, unsigned char* d_valid, void* stream);
/* Training targets from converged densities: d_pinfo [P][5] = (molecule, AO
 * offset of a, AO count of a, AO offset of b, AO count of b); d_Y, d_mask
 * [P][32] (block entries r*5+c, masked beyond the atoms' AO counts). */
int pscf_pair_targets(int64_t P, const int* d_pinfo, const unsigned char* d_valid, const void* d_mols,
                      const double* d_D, float* d_Y, float* d_mask, void* stream);
/* Write predicted blocks d_pred [P][32] into d_D (caller zeroes d_D). */
int pscf_pair_assemble(int64_t P, const int* d_pinfo, const unsigned char* d_valid, const void* d_mols,
                       const float* d_pred, double* d_D, void* stream);
/* BF16 MFMA GEMM with fused epilogues (row-major, FP32 accumulation):
 *   op 0: out = A B + bias (FP32), out2 = silu(out) (BF16)       [hidden layer]
 *   op 1: out = A B + bias (FP32)                                 [output layer]
 *   op 2: out = (A B^T) * silu'(aux) (FP32 and BF16 out2)         [backward data]
 *   op 3: out += A^T B (FP32 atomics; ksplit slices of K)         [weight gradient] */
int pscf_mlp_gemm(int op, int M, int N, int K, const void* d_A, int lda, const void* d_B, int ldb,
                  const float* d_bias, const float* d_aux, float* d_out, void* d_out2, int ldc, int ksplit,
                  void* stream);
/* Masked mean-squared-error gradient on [B][32] outputs: g = 2 (pred - Y) mask
 * inv_n (FP32 and BF16); the loss sum is added to *d_loss. */
int pscf_mlp_mse_grad(int B, const float* d_pred, const float* d_Y, const float* d_mask, float inv_n, float* d_g32,
                      void* d_g16, double* d_loss, void* stream);
/* Bias gradient: d_db[n] = sum_m dZ[m][n]. */
int pscf_mlp_colsum(int M, int N, const float* d_dz, int ld, float* d_db, void* stream);
/* Adam (bias-corrected, step >= 1) on FP32 master weights; refreshes the BF16
 * copy d_w16 when given. */
int pscf_mlp_adam(int64_t n, float* d_w, void* d_w16, const float* d_g, float* d_m, float* d_v, float lr, float b1,
                  float b2, float eps, int step, void* stream);
/* FP32 -> BF16 (round to nearest even). */
int pscf_mlp_to_bf16(int64_t n, const float* d_x, void* d_y, void* stream);

/* pscf_cast: elementwise FP64 <-> FP32 conversion helper (n elements). */
int pscf_cast(int src_dtype, const void* d_src, int dst_dtype, void* d_dst, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
