"""Electron repulsion integrals on the MI355X (drop-in for pyscf_ipu.electron_repulsion.direct).

Kept signatures (pyscf_ipu/electron_repulsion/direct.py):
  prepare_int_floats(mol)                                   :710-739
  prepare_integrals_2_inputs(mol) == prepare_electron_repulsion_integrals  :741-946
  electron_repulsion_integrals(input_floats, input_ints, mol, num_threads=3, v=1)  :513-515
  prepare_einsum_inputs(mol)                                :1029-1110
  ipu_einsum(__out3, dm, mol, threads=1, v=2) -> (vj, vk)   :500-504
plus the packed-s8 fast path used by the SCF driver (`eri_s8`).

The integer work list of `prepare_integrals_2_inputs` is reproduced exactly,
including the reference's shell-quartet filter i>=j, k>=l, i*j>=k*l (:772),
its grouping by block size with `np.argsort` (:914-933) and the per-class
scratch sizes of `get_shapes` (:813-911); the Python quadruple loop is
vectorised.  The integrals themselves run in the library's Rys kernel
(`pscf_eri_quartets` / `pscf_eri_s8`).
"""
from __future__ import annotations

import ctypes
from typing import List, Tuple

import numpy as np
import torch

from .. import _lib, device

BAS_SLOTS, NPRIM_OF, NCTR_OF, ANG_OF = 8, 2, 3, 1


def _ao_loc(bas: np.ndarray) -> np.ndarray:
    return np.cumsum(np.concatenate([np.zeros(1), (bas[:, 1] * 2 + 1) * bas[:, 3]])).astype(np.int32)


def prepare_int_floats(mol):
    """(input_floats (1, n_env) f64, input_ints (1, 6 + n_ao_loc + 6 natm + 8 nbas) i32)."""
    atm, bas, env = mol._atm, mol._bas, mol._env
    n_atm, n_bas = atm.shape[0], bas.shape[0]
    ao_loc = _ao_loc(bas)
    n_ao_loc = int(np.prod(ao_loc.shape))
    n_buf, n_eri, n_env = 81, 81, int(np.prod(env.shape))
    input_floats = env.reshape(1, -1)
    input_ints = np.zeros((1, 6 + n_ao_loc + n_atm * 6 + n_bas * 8), dtype=np.int32)
    s = 0
    input_ints[:, s:s + 6] = np.array([n_eri, n_buf, n_atm, n_bas, n_env, n_ao_loc])
    s += 6
    input_ints[:, s:s + n_ao_loc] = ao_loc.reshape(-1)
    s += n_ao_loc
    input_ints[:, s:s + n_atm * 6] = atm.reshape(-1)
    s += n_atm * 6
    input_ints[:, s:s + n_bas * 8] = bas.reshape(-1)
    return input_floats, input_ints


def quartet_list(n_bas: int) -> np.ndarray:
    """(num_calls, 4) int32 (i, j, k, l) in the reference's loop order with its
    filter i >= j and k >= l and i*j >= k*l (direct.py:772-787)."""
    r = np.arange(n_bas, dtype=np.int64)
    i, j = np.meshgrid(r, r, indexing="ij")
    m = i >= j
    i, j = i[m], j[m]  # pairs with i >= j, row-major over (i, j)
    # For each (i, j) in loop order, all (k, l) with k >= l in loop order and i*j >= k*l.
    k, l = np.meshgrid(r, r, indexing="ij")
    mk = k >= l
    k, l = k[mk], l[mk]
    kl_prod = k * l
    # Lexicographic (i, j, k, l) order: (i, j) blocks each followed by their (k, l) list;
    # the reference's j loop runs over all j for each i and skips j > i, same for l.
    ij_prod = i * j
    sel = kl_prod[None, :] <= ij_prod[:, None]
    ii = np.repeat(i, sel.sum(axis=1))
    jj = np.repeat(j, sel.sum(axis=1))
    kk = np.broadcast_to(k, sel.shape)[sel]
    ll = np.broadcast_to(l, sel.shape)[sel]
    return np.stack([ii, jj, kk, ll], axis=1).astype(np.int32)


def _get_shapes(bas: np.ndarray, ijkl) -> Tuple[float, float, float]:
    """Per-class scratch sizes (glen, nf, buflen) of the reference's `get_shapes`
    (direct.py:813-911), reproduced for interface parity."""
    i_sh, j_sh, k_sh, l_sh = ijkl
    flat = bas.reshape(-1)
    i_ctr, j_ctr, k_ctr, l_ctr = (flat[BAS_SLOTS * s + NCTR_OF] for s in (i_sh, j_sh, k_sh, l_sh))
    i_l, j_l, k_l, l_l = (flat[BAS_SLOTS * s + ANG_OF] for s in (i_sh, j_sh, k_sh, l_sh))
    nfi, nfj, nfk, nfl = ((x + 1) * (x + 2) / 2 for x in (i_l, j_l, k_l, l_l))
    nf = nfi * nfk * nfl * nfj
    nc = i_ctr * j_ctr * k_ctr * l_ctr
    lenl = nf * nc
    lenk = nf * i_ctr * j_ctr * k_ctr
    lenj = nf * i_ctr * j_ctr
    leni = nf * i_ctr
    len0 = nf
    nrys_roots = (i_l + j_l + k_l + l_l) / 2 + 1
    ibase = i_l > j_l
    kbase = k_l > l_l
    if nrys_roots <= 2:
        ibase = 0
        kbase = 0
    if kbase:
        dlk, dll = k_l + l_l + 1, l_l + 1
    else:
        dlk, dll = k_l + 1, k_l + l_l + 1
    if ibase:
        dli, dlj = i_l + j_l + 1, j_l + 1
    else:
        dli, dlj = i_l + 1, i_l + j_l + 1
    g_size = nrys_roots * dli * dlk * dll * dlj
    leng = g_size * 3 * ((1 << 0) + 1)
    length = leng + lenl + lenk + lenj + leni + len0
    nfik = nfi * nfk
    nfikl = nfik * nfl
    dj = j_l * 2 + 1
    buflen = nfikl * dj
    return length, nf, buflen


def prepare_integrals_2_inputs(mol):
    """Same 9-tuple as the reference: (input_floats, input_ints, tuple_ijkl,
    shapes, sizes, counts, indxs, indxs_inv, num_calls)."""
    bas = mol._bas
    ao_loc = _ao_loc(bas)
    q = quartet_list(bas.shape[0])
    num_calls = len(q)
    d = (ao_loc[q + 1] - ao_loc[q]).astype(np.float64)  # (num_calls, 4) di, dj, dk, dl
    output_sizes = np.concatenate([d, np.prod(d, axis=1, keepdims=True)], axis=1)
    input_floats, input_ints = prepare_int_floats(mol)
    env32 = input_floats.astype(np.float32)  # the reference hands float32 env to the device
    indxs = np.argsort(output_sizes[:, -1])
    input_ijkl = q[indxs]
    sizes, counts = np.unique(output_sizes[:, -1], return_counts=True)
    sizes, counts = sizes.astype(np.int32), counts.astype(np.int32)
    inputs, start = [], 0
    for size, count in zip(sizes, counts):
        inputs.append(tuple(map(tuple, input_ijkl[start:start + count].tolist())))
        start += count
    tuple_ijkl = tuple(inputs)
    shapes = tuple(_get_shapes(bas, grp[0]) for grp in inputs)
    indxs_inv = np.arange(indxs.shape[0])
    indxs_inv[indxs] = np.arange(indxs.shape[0])
    return (env32, input_ints, tuple_ijkl, shapes, tuple(sizes.tolist()), tuple(counts.tolist()), indxs, indxs_inv,
            num_calls)


prepare_electron_repulsion_integrals = prepare_integrals_2_inputs


def _mol_device(mol):
    atm = device.to_dev(mol._atm.astype(np.int32))
    bas = device.to_dev(mol._bas.astype(np.int32))
    env = device.to_dev(mol._env.astype(np.float64))
    return atm, bas, env


def eri_quartets(mol, quartets: np.ndarray, dtype: str = "fp32") -> torch.Tensor:
    """Blocks of an explicit quartet list, concatenated in list order (flat)."""
    tdt, code, _ = device.DTYPES[dtype]
    atm, bas, env = _mol_device(mol)
    ao_loc = _ao_loc(mol._bas)
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    sz = np.prod((ao_loc[q + 1] - ao_loc[q]).astype(np.int64), axis=1)
    offs = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64)
    out = device.empty(int(sz.sum()), tdt)
    dq, do = device.to_dev(q), device.to_dev(offs)
    _lib.call("pscf_eri_quartets", code, device.ptr(atm), len(mol._atm), device.ptr(bas), len(mol._bas),
              device.ptr(env), device.ptr(dq), device.ptr(do), len(q), device.ptr(out), device.stream())
    return out


def electron_repulsion_integrals(input_floats, input_ints, mol, num_threads: int = 3, v: int = 1):
    """List over block-size classes of (count, size) FP32 device arrays, rows in
    the order of `prepare_electron_repulsion_integrals(mol)[2]`, each block in
    libcint's (dl, dk, dj, di) element order.  `input_floats`/`input_ints` are
    accepted for signature compatibility; the molecule's own arrays are used."""
    _, _, tuple_ijkl, _, sizes, counts, _, _, _ = prepare_electron_repulsion_integrals(mol)
    flat = eri_quartets(mol, np.concatenate([np.asarray(t, dtype=np.int32).reshape(-1, 4) for t in tuple_ijkl]),
                        "fp32")
    out, s = [], 0
    for size, count in zip(sizes, counts):
        out.append(flat[s:s + size * count].view(count, size))
        s += size * count
    return out


def prepare_einsum_inputs(mol):
    """(tuple_indices, tuple_do_lists, N, num_calls) as the reference
    (direct.py:1029-1110): per quartet of `quartet_list` (loop order)
    (di, dj, dk, dl, i0, j0, k0, l0) and the 8 flags of the index images of the
    block origin that this quartet applies (the first quartet, in loop order,
    and the first image within it, that produces an origin applies it)."""
    ind, do = _einsum_tables(mol._bas)
    return tuple(map(tuple, ind.tolist())), tuple(map(tuple, do.tolist())), mol.nao_nr(), len(ind)


def _einsum_tables(bas: np.ndarray):
    ao_loc = _ao_loc(bas).astype(np.int64)
    q = quartet_list(bas.shape[0]).astype(np.int64)
    d = ao_loc[q + 1] - ao_loc[q]
    o = ao_loc[q]
    i0, j0, k0, l0 = o[:, 0], o[:, 1], o[:, 2], o[:, 3]
    images = [(i0, j0, k0, l0), (i0, j0, l0, k0), (j0, i0, k0, l0), (j0, i0, l0, k0),
              (k0, l0, i0, j0), (k0, l0, j0, i0), (l0, k0, i0, j0), (l0, k0, j0, i0)]
    # origins < 2^15: one int64 key per image; flattened in (quartet, image) order
    keys = np.stack([(a << 45) | (b << 30) | (c << 15) | e for a, b, c, e in images], axis=1).reshape(-1)
    _, first = np.unique(keys, return_index=True)
    do = np.zeros(keys.shape[0], np.int32)
    do[first] = 1
    return np.concatenate([d, o], axis=1).astype(np.int32), do.reshape(-1, 8)


def ipu_einsum(__out3, dm, mol, threads: int = 1, v: int = 2):
    """(vj, vk) with vj[k, l] = sum (ij|kl) dm[j, i] and vk[i, l] = sum (ij|kl)
    dm[j, k] (the einsums 'ijkl,ji->kl' / 'ijkl,jk->il'), from the ERI blocks of
    `electron_repulsion_integrals` (one (count, size) array per block size, in
    the order of prepare_electron_repulsion_integrals), block-wise in one launch
    (`pscf_jk_blocks`).  `threads` and `v` select IPU tile variants in the
    reference and are accepted for signature compatibility.  Arithmetic in the
    ERI blocks' dtype; returns device tensors, or numpy arrays for a numpy dm."""
    is_np = not isinstance(dm, torch.Tensor)
    N = mol.nao_nr()
    if N > 0xffff:
        raise ValueError("ipu_einsum: nao above 65535")
    ind, do = _einsum_tables(mol._bas)
    *_, indxs, _, _ = prepare_electron_repulsion_integrals(mol)
    ind, do = ind[indxs], do[indxs]
    sizes = np.prod(ind[:, :4].astype(np.int64), axis=1)
    parts = [p if isinstance(p, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(p)) for p in __out3]
    dt = torch.float64 if parts[0].dtype == torch.float64 else torch.float32
    code = _lib.F64 if dt == torch.float64 else _lib.F32
    counts = [int(p.shape[0]) for p in parts]
    if sum(counts) != len(ind):
        raise ValueError(f"ipu_einsum: {sum(counts)} ERI blocks given, the molecule has {len(ind)}")
    cls_size, cls_blk0, cls_elem0, b0, e0 = [], [], [0], 0, 0
    for p, c in zip(parts, counts):
        if c == 0:
            continue
        sz = int(sizes[b0])
        if not np.all(sizes[b0:b0 + c] == sz) or p.numel() != c * sz:
            raise ValueError("ipu_einsum: ERI blocks do not follow the molecule's block-size classes")
        cls_size.append(sz)
        cls_blk0.append(b0)
        b0 += c
        e0 += c * sz
        cls_elem0.append(e0)
    eri = torch.cat([device.to_dev(p.reshape(-1), dt) for p in parts if p.shape[0] > 0])
    mask = (do << np.arange(8)).sum(axis=1).astype(np.int64)
    d = ind[:, :4].astype(np.int64)
    o = ind[:, 4:].astype(np.int64)
    tab = np.zeros((len(ind), 4), np.int64)
    tab[:, 0] = d[:, 0] | (d[:, 1] << 4) | (d[:, 2] << 8) | (d[:, 3] << 12) | (mask << 16)
    tab[:, 1] = o[:, 0] | (o[:, 1] << 16)
    tab[:, 2] = o[:, 2] | (o[:, 3] << 16)
    tab = device.to_dev(tab.astype(np.uint32).view(np.int32))
    dmt = device.to_dev(dm if not is_np else np.asarray(dm), dt).reshape(N, N).contiguous()
    vj, vk = device.empty((N, N), dt), device.empty((N, N), dt)
    nc = len(cls_size)
    wsb = _lib.lib().pscf_jk_blocks_workspace_bytes(code, N, e0)
    ws = device.workspace(wsb)
    _lib.call("pscf_jk_blocks", code, device.ptr(eri), device.ptr(tab), nc,
              (ctypes.c_int64 * (nc + 1))(*cls_elem0), (ctypes.c_int * nc)(*cls_blk0),
              (ctypes.c_int * nc)(*cls_size), device.ptr(dmt), N, device.ptr(vj), device.ptr(vk), device.ptr(ws),
              wsb, device.stream())
    if is_np:
        return vj.cpu().numpy(), vk.cpu().numpy()
    return vj, vk


def shell_pairs_by_class(mol):
    """Shell pairs (i, j) of all i >= j, grouped ss | ps (p shell first) | pp.

    Returns (pairs (npair, 2) int32, counts [n_ss, n_ps, n_pp])."""
    l = np.asarray(mol._bas[:, ANG_OF])
    i, j = np.tril_indices(len(l))            # (i, j), i >= j, in row order
    first = l[i] >= l[j]
    ab = np.stack([np.where(first, i, j), np.where(first, j, i)], axis=1).astype(np.int32)
    cls = l[i] + l[j]
    order = np.argsort(cls, kind="stable")   # ss, then ps, then pp, each in row order
    return np.ascontiguousarray(ab[order]), [int(np.sum(cls == c)) for c in range(3)]


def s8_size(nao: int) -> int:
    P = nao * (nao + 1) // 2
    return P * (P + 1) // 2


def eri_s8(mol, dtype: str = "fp64", out: torch.Tensor = None, dev_mol=None, dev_pairs=None) -> torch.Tensor:
    """All unique ERIs in PySCF's aosym='s8' packing, on the device.
    `dev_pairs` = (device pairs, counts) of shell_pairs_by_class, if uploaded."""
    tdt, code, _ = device.DTYPES[dtype]
    if any(mol._bas[:, ANG_OF] > 1):
        raise _lib.HIPError("eri_s8: only s and p shells are supported")
    atm, bas, env = dev_mol if dev_mol is not None else _mol_device(mol)
    nao = mol.nao_nr()
    if out is None:
        out = device.empty(s8_size(nao), tdt)
    if dev_pairs is None:
        pairs, counts = shell_pairs_by_class(mol)
        dp = device.to_dev(pairs)
    else:
        dp, counts = dev_pairs
    cc = (ctypes.c_int * 3)(*counts)
    _lib.call("pscf_eri_s8", code, device.ptr(atm), len(mol._atm), device.ptr(bas), len(mol._bas), device.ptr(env),
              device.ptr(dp), cc, nao, device.ptr(out), device.stream())
    return out
