"""Drop-in replacement of impurityModel.ed.spectra's Green's functions on the GPU.

``getGreen``, ``getSpectra``, ``getRIXSmap`` and ``simulate_spectra`` keep
the reference's signatures (impurityModel/ed/spectra.py:36-1130).
``getSpectra`` runs all (transition operator, eigenstate) chains of a call
as one batch: one BFS closure launch sequence for all chains, one Lanczos
launch and one continued-fraction launch.  Its T|psi_i> follow the
reference's memo semantics for the chosen ``parallelization_mode``.

The sections of ``simulate_spectra`` are library calls of their own:
``ps_spectra``, ``xps_spectra``, ``nixs_spectra``, ``xas_spectra`` and
``rixs_spectra`` (spectra.py:124-293), each returning the Green's
functions, their thermal averages and the operator sum.
"""

from math import sqrt

import numpy as np

from . import opcompile
from .engine import Batch, Ops, vec_apply, vec_from_host
from .finite import _ops_for, _words
from .average import thermal_average
from .models import K_B  # noqa: F401  (k_B in eV/K)
from .operators import (  # noqa: F401  (reference names)
    getDaggeredDipoleOperators,
    getDipoleOperator,
    getDipoleOperators,
    getInversePhotoEmissionOperators,
    getNIXSOperator,
    getNIXSOperators,
    getPhotoEmissionOperators,
)


def getGreen(n_spin_orbitals, e, psi, hOp, omega, delta, krylovSize, slaterWeightMin, restrictions=None,
             h_dict=None, mode="sparse", parallelization_mode="serial", reorth="none"):
    """<psi|((omega + i delta + e) - H)^-1|psi> (spectra.py:551-681) on the GPU.

    BFS closure of psi's states, CSR, Lanczos (K = min(krylovSize, n)) and
    the continued fraction all run on the device.
    """
    omega = np.asarray(omega, dtype=np.float64)
    if len(psi) == 0:
        return np.zeros(len(omega), dtype=complex)
    if mode not in ("sparse", "dense", "dict"):
        raise Exception("Value of variable 'mode' is incorrect.")
    n = n_spin_orbitals
    keys = list(psi.keys())
    amps = np.array([complex(x) for x in psi.values()])
    ops = _ops_for(n, hOp, restrictions)
    b = Batch(ops, [0])
    b.set_seeds([_words(keys, n)], [amps])
    b.build(fixed_basis=False, weight_min=1e-12)
    K = int(min(krylovSize, b.chain_n[0]))
    _, _, kout, status = b.lanczos(K, reorth=1 if reorth == "full" else 0, want_coeffs=False)
    if status[0] != 0:
        raise ValueError("Warning: beta==0, implementation absent!")
    g, _ = b.cf([e], [1.0], [1], [1], 1.0, [0], 1, omega, delta, want_g=True)
    return g[0]


def _transition_states_memo(n, tOps, psis, slaterWeightMin, restrictions):
    """T_t|psi_i> for chains (t, i) with one ``t_big`` memo per operator
    shared over the eigenstates (spectra.py:779-786): a product state met by
    an earlier eigenstate reuses its stored, pruned T|state>
    (finite.py:1413-1504).  Returns per-chain (keys [n_c, W], amplitudes)."""
    from .finite import applyOp

    W = opcompile.n_words(n)
    keys, amps = [], []
    for tOp in tOps:
        t_big = {}
        for psi in psis:
            psiR = applyOp(n, tOp, psi, slaterWeightMin, restrictions, t_big)
            ks = list(psiR.keys())
            keys.append(_words(ks, n) if ks else np.zeros((0, W), np.uint64))
            amps.append(np.array([complex(x) for x in psiR.values()], np.complex128))
    return keys, amps


def getSpectra(n_spin_orbitals, hOp, tOps, psis, es, w, delta, restrictions=None, krylovSize=150,
               slaterWeightMin=1e-7, parallelization_mode="H_build", reorth="none", info=None):
    """gs[i, t, :] = |T_t psi_i|^2 <T_t psi_i|(w + i delta + e_i - H)^-1|T_t psi_i> / |T_t psi_i|^2.

    spectra.py:684-806.  A negative ``delta`` with ``-w`` (photoemission) is
    handled exactly as the reference: z = w + i delta + e with the values
    given.  ``parallelization_mode`` selects the reference's two ways of
    forming T|psi_i>: "H_build" (its default) keeps one memo per operator
    over the eigenstates, whose stored per-state results are pruned at
    ``slaterWeightMin`` after their first use; "eigen_states" applies T to
    every eigenstate afresh (one batched application).  Either way every
    chain's BFS closure, Lanczos and continued fraction run in one batch.
    An empty T|psi_i> gives a zero Green's function (spectra.py:611-612).
    ``info`` (a dict) receives ``chain_n``: the BFS basis size of every
    non-empty chain in (operator, eigenstate) order, as the reference logs it.
    """
    if parallelization_mode not in ("H_build", "eigen_states"):
        raise Exception("Incorrect value of variable parallelization_mode.")
    w = np.asarray(w, dtype=np.float64)
    ne, nt = len(es), len(tOps)
    n = n_spin_orbitals
    gs = np.zeros((ne, nt, len(w)), dtype=complex)
    if ne == 0 or nt == 0:
        return gs
    W = opcompile.n_words(n)
    # chain o = (t, i) in the reference's loop order
    out_src = np.array([i for t in range(nt) for i in range(ne)], np.int32)
    out_model = np.array([t for t in range(nt) for i in range(ne)], np.int32)
    if parallelization_mode == "H_build":
        keys, amps = _transition_states_memo(n, tOps, psis, slaterWeightMin, restrictions)
    else:
        tops = Ops(list(tOps), [restrictions] * nt, n, W)
        src = vec_from_host(W, [_words(list(p.keys()), n) for p in psis],
                            [np.array([complex(x) for x in p.values()]) for p in psis])
        res = vec_apply(tops, src, out_src, out_model, slaterWeightMin)
        k, a, _ = res.get()
        off = res.offsets
        keys = [k[off[o]:off[o + 1]] for o in range(len(out_src))]
        amps = [a[off[o]:off[o + 1]] for o in range(len(out_src))]
    norm2 = np.array([float(np.sum(np.abs(x) ** 2)) for x in amps])
    live = np.nonzero(norm2 > 0)[0]
    if len(live) == 0:
        return gs
    seeds = vec_from_host(W, [keys[o] for o in live], [amps[o] / np.sqrt(norm2[o]) for o in live])
    hops = _ops_for(n, hOp, restrictions)
    b = Batch(hops, np.zeros(len(live), np.int32))
    b.seed_from_vec(seeds)
    b.build(fixed_basis=False, weight_min=1e-12)
    if info is not None:
        info.setdefault("chain_n", []).extend(int(x) for x in b.chain_n)
    # Krylov size per chain is min(krylovSize, n_c) (spectra.py:672); chains
    # shorter than K stop at their dimension inside the kernel.
    K = int(min(krylovSize, max(int(b.chain_n.max()), 1)))
    _, _, kout, status = b.lanczos(K, reorth=1 if reorth == "full" else 0, want_coeffs=False)
    if np.any(status != 0):
        raise ValueError("Warning: beta==0, implementation absent!")
    energy = np.asarray(es, np.float64)[out_src[live]]
    # sign = +1: z = w + i delta + e with the caller's (possibly negated) w, delta
    # one spectrum id per chain: one thread per (chain, omega)
    ones = np.ones(len(live), np.int32)
    g, _ = b.cf(energy, norm2[live], ones, ones, 1.0, np.arange(len(live), dtype=np.int32), len(live), w, delta,
                want_g=True)
    for j, o in enumerate(live):
        gs[out_src[o], out_model[o], :] = g[j]
    return gs


def _rixs_chains(n, hops, tin, tout, src, es, nin, nout, wIns, wLoss, delta1, delta2, krylovSize, slaterWeightMin,
                 alias_memo, reorth, solve_tol, solve_max_steps):
    """All RIXS chains of one model from eigenstates ``src`` (device vectors).

    Returns ``g [ne, nin, nout, nwIn, nwLoss]`` and a timing dict.  Every
    stage is one batched device call over all chains of the map.
    """
    import time

    tim = {}
    ne, nwi = len(es), len(wIns)
    t0 = time.perf_counter()
    # psi1 = T_in psi for chain (tIn, iE) in the reference's loop order
    c_src = np.array([i for t in range(nin) for i in range(ne)], np.int32)
    c_mod = np.array([t for t in range(nin) for i in range(ne)], np.int32)
    psi1 = vec_apply(tin, src, c_src, c_mod, slaterWeightMin, normalize=False)
    eb = Batch(hops, np.zeros(len(c_src), np.int32))
    eb.seed_from_vec(psi1)
    eb.build(fixed_basis=False, weight_min=1e-12)
    tim["core_n"] = eb.chain_n.copy()
    tim["core_hole_build"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    e_c = np.asarray(es, np.float64)[c_src]
    shifts = wIns[None, :] + 1j * delta1 + e_c[:, None]
    _, resid, ok = eb.solve(shifts, tol=solve_tol, max_steps=solve_max_steps, allow_noconv=True)
    if not ok:
        # the reference reports bicgstab's non-convergence and carries on
        print("Rank 0: Convergence to tolerance not achieved")
        print("max relative residual = ", float(np.nanmax(resid)))
    C = len(c_src)
    psi2 = eb.vec_from_eigen(np.repeat(np.arange(C, dtype=np.int32), nwi), np.tile(np.arange(nwi, dtype=np.int32), C),
                             0.0)
    tim["solve"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    # psi3 = T_out psi2 for chain ((tIn, iE), iwIn, tOut), normalized, pruned
    V2 = C * nwi
    o_src = np.repeat(np.arange(V2, dtype=np.int32), nout)
    o_mod = np.zeros(V2 * nout, np.int32) if alias_memo else np.tile(np.arange(nout, dtype=np.int32), V2)
    psi3 = vec_apply(tout, psi2, o_src, o_mod, slaterWeightMin, normalize=True)
    norm2 = psi3.norm2()
    psi3 = psi3.prune(slaterWeightMin)
    gb = Batch(hops, np.zeros(len(o_src), np.int32))
    gb.seed_from_vec(psi3)
    gb.build(fixed_basis=False, weight_min=1e-12)
    tim["out_n"] = gb.chain_n.copy()
    tim["ground_build"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    K = int(min(krylovSize, max(int(gb.chain_n.max()), 1)))
    _, _, _, status = gb.lanczos(K, reorth=1 if reorth == "full" else 0, want_coeffs=False)
    if np.any(status != 0):
        raise ValueError("Warning: beta==0, implementation absent!")
    energy = np.repeat(e_c, nwi * nout)
    ones = np.ones(len(o_src), np.int32)
    # one spectrum id per chain: one thread per (chain, loss energy)
    g, _ = gb.cf(energy, norm2, ones, ones, 1.0, np.arange(len(o_src), dtype=np.int32), len(o_src), wLoss, delta2,
                 want_g=True)
    tim["lanczos_cf"] = time.perf_counter() - t0
    # chain index = ((tIn * ne + iE) * nwi + iwIn) * nout + tOut
    g = g.reshape(nin, ne, nwi, nout, len(wLoss)).transpose(1, 0, 3, 2, 4)
    return g, tim


def getRIXSmap(n_spin_orbitals, hOp, tOpsIn, tOpsOut, psis, es, wIns, wLoss, delta1, delta2, restrictions=None,
               krylovSize=150, slaterWeightMin=1e-7, h_dict_ground=None, parallelization_mode="H_build_wIn",
               alias_memo=False, reorth="none", solve_tol=1e-12, solve_max_steps=1000, info=None):
    """RIXS Green's functions gs[iE, tIn, tOut, iwIn, iwLoss] (spectra.py:809-1130).

    g = |psi3|^2 <psi3|(wLoss + i delta2 + e - H)^-1|psi3>, psi3 the
    normalized T_out (wIn + i delta1 + e - H)^-1 T_in psi, pruned at
    slaterWeightMin.  The resolvent solves of all incoming energies of a
    (T_in, eigenstate) pair share one Krylov space on the core-hole sector
    (relative residual <= solve_tol, where the reference's bicgstab stops at
    scipy's default 1e-5); every outgoing chain is one chain of a single
    batched BFS + Lanczos + continued-fraction pass.

    ``alias_memo=True`` reproduces the reference's one memo shared by all
    out-going operators (``[{}] * len(tOpsOut)``, spectra.py:910), under
    which every out-going channel uses the first operator.  ``h_dict_ground``
    and ``parallelization_mode`` are accepted; the memo is not needed on the
    device and an unknown mode returns zeros as the reference does.
    ``info`` (a dict) receives the stage times and the BFS sizes of the
    core-hole sectors (``core_n``, per (T_in, eigenstate)) and of the
    out-going chains (``out_n``, per (T_in, eigenstate, wIn, T_out)).
    """
    wIns = np.asarray(wIns, dtype=np.float64)
    wLoss = np.asarray(wLoss, dtype=np.float64)
    ne, nin, nout = len(es), len(tOpsIn), len(tOpsOut)
    gs = np.zeros((ne, nin, nout, len(wIns), len(wLoss)), dtype=complex)
    if parallelization_mode not in ("serial", "H_build", "wIn", "H_build_wIn"):
        return gs
    if ne == 0 or nin == 0 or nout == 0 or len(wIns) == 0:
        return gs
    n = n_spin_orbitals
    W = opcompile.n_words(n)
    hops = _ops_for(n, hOp, restrictions)
    tin = Ops(list(tOpsIn), [restrictions] * nin, n, W)
    tout = Ops(list(tOpsOut), [restrictions] * nout, n, W)
    src = vec_from_host(W, [_words(list(p.keys()), n) for p in psis],
                        [np.array([complex(x) for x in p.values()]) for p in psis])
    g, tim = _rixs_chains(n, hops, tin, tout, src, es, nin, nout, wIns, wLoss, delta1, delta2, krylovSize,
                          slaterWeightMin, alias_memo, reorth, solve_tol, solve_max_steps)
    if info is not None:
        info.update(tim)
    gs[:] = g
    return gs


# ------------------------------------------------------------------ simulate_spectra compositions


def _thermal(es, gs, T):
    """Thermal average of -Im gs over its eigenstate axis (average.py:11-45)."""
    return thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)


def ps_spectra(n_spin_orbitals, hOp, psis, es, w, delta, restrictions, nBaths, T=300, krylovSize=150,
               slaterWeightMin=1e-7, l=2, reorth="none", info=None):
    """3d inverse photoemission + photoemission (spectra.py:124-149).

    gs = -getSpectra(c_i, -w, -delta) + getSpectra(c_i^dagger, w, delta),
    one entry per spin-orbital of shell ``l``.  Returns ``(gs, a, aSum)``:
    the Green's functions [n_eig, n_ops, n_w], their thermal average of -Im
    per operator and its sum over operators.
    """
    w = np.asarray(w, np.float64)
    tOpsIPS = getInversePhotoEmissionOperators(nBaths, l=l)
    tOpsPS = getPhotoEmissionOperators(nBaths, l=l)
    gsIPS = getSpectra(n_spin_orbitals, hOp, tOpsIPS, psis, es, w, delta, restrictions, krylovSize, slaterWeightMin,
                       reorth=reorth, info=info)
    gsPS = getSpectra(n_spin_orbitals, hOp, tOpsPS, psis, es, -w, -delta, restrictions, krylovSize, slaterWeightMin,
                      reorth=reorth, info=info)
    gs = -gsPS + gsIPS
    a = _thermal(es, gs, T)
    return gs, a, np.sum(a, axis=0)


def xps_spectra(n_spin_orbitals, hOp, psis, es, w, delta, restrictions, nBaths, T=300, krylovSize=150,
                slaterWeightMin=1e-7, reorth="none", info=None):
    """Core 2p x-ray photoemission (spectra.py:159-179): gs = -getSpectra(c_2p, -w, -delta)."""
    w = np.asarray(w, np.float64)
    tOpsPS = getPhotoEmissionOperators(nBaths, l=1)
    gs = -getSpectra(n_spin_orbitals, hOp, tOpsPS, psis, es, -w, -delta, restrictions, krylovSize, slaterWeightMin,
                     reorth=reorth, info=info)
    a = _thermal(es, gs, T)
    return gs, a, np.sum(a, axis=0)


def nixs_spectra(n_spin_orbitals, hOp, psis, es, wLoss, deltaNIXS, qsNIXS, liNIXS, ljNIXS, RiNIXS, RjNIXS, radialMesh,
                 restrictions, nBaths, T=300, krylovSize=150, slaterWeightMin=1e-7, reorth="none", info=None):
    """Non-resonant inelastic x-ray scattering (spectra.py:188-208): one
    transition operator exp(i q.r) per momentum of ``qsNIXS`` on ``wLoss``."""
    tOps = getNIXSOperators(nBaths, qsNIXS, liNIXS, ljNIXS, RiNIXS, RjNIXS, radialMesh)
    gs = getSpectra(n_spin_orbitals, hOp, tOps, psis, es, wLoss, deltaNIXS, restrictions, krylovSize, slaterWeightMin,
                    reorth=reorth, info=info)
    a = _thermal(es, gs, T)
    return gs, a, np.sum(a, axis=0)


def xas_spectra(n_spin_orbitals, hOp, psis, es, w, delta, epsilons, restrictions, nBaths, T=300, krylovSize=150,
                slaterWeightMin=1e-7, reorth="none", info=None):
    """X-ray absorption (spectra.py:218-238): one 2p -> 3d dipole operator per
    polarization of ``epsilons``."""
    tOps = getDipoleOperators(nBaths, epsilons)
    gs = getSpectra(n_spin_orbitals, hOp, tOps, psis, es, w, delta, restrictions, krylovSize, slaterWeightMin,
                    reorth=reorth, info=info)
    a = _thermal(es, gs, T)
    return gs, a, np.sum(a, axis=0)


def rixs_spectra(n_spin_orbitals, hOp, psis, es, wIn, wLoss, delta, deltaRIXS, epsilonsRIXSin, epsilonsRIXSout,
                 restrictions, nBaths, T=300, krylovSize=150, slaterWeightMin=1e-7, alias_memo=True,
                 reorth="none", info=None):
    """2p -> 3d -> 2p RIXS (spectra.py:248-293).

    ``alias_memo=True`` (the default here) is the reference's behaviour: its
    out-going memo list ``[{}] * len(tOpsOut)`` (spectra.py:910) shares one
    dict, so every out-going channel uses the first out-going operator.
    Returns ``(gs, a, aSum)`` with gs [n_eig, n_in, n_out, n_wIn, n_wLoss],
    a the thermal average of -Im gs and aSum its sum over polarizations
    [n_wIn, n_wLoss].
    """
    tOpsIn = getDipoleOperators(nBaths, epsilonsRIXSin)
    tOpsOut = getDaggeredDipoleOperators(nBaths, epsilonsRIXSout)
    gs = getRIXSmap(n_spin_orbitals, hOp, tOpsIn, tOpsOut, psis, es, wIn, wLoss, delta, deltaRIXS, restrictions,
                    krylovSize, slaterWeightMin, alias_memo=alias_memo, reorth=reorth, info=info)
    a = _thermal(es, gs, T)
    return gs, a, np.sum(a, axis=(0, 1))


def simulate_spectra(es, psis, hOp, T, w, delta, epsilons, wLoss, deltaNIXS, qsNIXS, liNIXS, ljNIXS, RiNIXS, RjNIXS,
                     radialMesh, wIn, deltaRIXS, epsilonsRIXSin, epsilonsRIXSout, restrictions, h5f, nBaths,
                     output_dir=".", write_files=True):
    """PS/IPS, XPS, NIXS, XAS and RIXS of the given eigenstates (spectra.py:36-299).

    The same arguments, sections, sign conventions and outputs as the
    reference: each section's Green's functions and thermal averages go to
    ``h5f`` (any object with ``create_dataset(name, data=...)``, or None),
    and ``PS.dat``, ``XPS.dat``, ``NIXS.dat``, ``XAS.dat`` and ``RIXS.bin``
    are written to ``output_dir`` unless ``write_files`` is False.  Returns
    a dict {section: (gs, a, aSum)} and, under "info", each section's chain
    sizes and stage times.
    """
    import os
    import time

    from .output import write_dat, write_rixs_bin

    n = sum(2 * (2 * ang + 1) + nb for ang, nb in nBaths.items())
    w = np.asarray(w, np.float64)
    wLoss = np.asarray(wLoss, np.float64)
    wIn = np.asarray(wIn, np.float64)
    out = {"info": {k: {} for k in ("PS", "XPS", "NIXS", "XAS", "RIXS")}}
    info = out["info"]

    def record(name, res, grid, timer):
        gs, a, aSum = res
        out[name] = res
        if h5f is not None:
            h5f.create_dataset(name, data=-gs.imag)
            h5f.create_dataset(name + "thermal", data=a)
        if write_files:
            if name == "RIXS":
                write_rixs_bin(os.path.join(output_dir, "RIXS.bin"), wIn, wLoss, aSum)
            else:
                write_dat(os.path.join(output_dir, name + ".dat"), grid, a)
        print("time({}) = {:.2f} seconds \n".format(name, time.time() - timer))

    t0 = time.time()
    record("PS", ps_spectra(n, hOp, psis, es, w, delta, restrictions, nBaths, T, info=info["PS"]), w, t0)
    t0 = time.time()
    record("XPS", xps_spectra(n, hOp, psis, es, w, delta, restrictions, nBaths, T, info=info["XPS"]), w, t0)
    t0 = time.time()
    record("NIXS", nixs_spectra(n, hOp, psis, es, wLoss, deltaNIXS, qsNIXS, liNIXS, ljNIXS, RiNIXS, RjNIXS, radialMesh,
                                restrictions, nBaths, T, info=info["NIXS"]), wLoss, t0)
    t0 = time.time()
    record("XAS", xas_spectra(n, hOp, psis, es, w, delta, epsilons, restrictions, nBaths, T, info=info["XAS"]), w, t0)
    t0 = time.time()
    record("RIXS", rixs_spectra(n, hOp, psis, es, wIn, wLoss, delta, deltaRIXS, epsilonsRIXSin, epsilonsRIXSout,
                                restrictions, nBaths, T, info=info["RIXS"]), None, t0)
    return out
