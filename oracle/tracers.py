"""CPU restatement of the E3SM tracer-array transfer (mam4xx utils.hpp:
183-426) -- test infrastructure: the checker of mam4x_tracers_to_fields and
mam4x_fields_to_tracers, never part of the product.

The reference walks a running index through the tracer row: gases from
index 9 (gasses_start_ind, utils.hpp:36-40), then per mode its species in
mode_aero_species order and its number (aero_start_ind = 15,
utils.hpp:44-48).  qqcw uses the same indices from 15.  The functions below
follow that walk with the reference's Prognostics shapes: q_gas[6] and, per
mode m, n_mode_{i,c}[m] and q_aero_{i,c}[m][a] with a the mode-local species
index.  Each entry is an array over cells; None is an absent view.
"""
import numpy as np

GASES_START = 9       # utils.hpp:36-40
AERO_START = 15       # utils.hpp:44-48
NUM_GAS = 6           # O3 H2O2 H2SO4 SO2 DMS SOAG
NSPEC_MODE = (7, 4, 7, 3)  # mam4::num_species_mode
PCNST = 40            # mam4_config.hpp.in:25


def empty_progs():
    return dict(q_gas=[None] * NUM_GAS,
                n_mode_i=[None] * 4, n_mode_c=[None] * 4,
                q_aero_i=[[None] * 7 for _ in range(4)],
                q_aero_c=[[None] * 7 for _ in range(4)])


def extract_stateq_from_prognostics(progs, q, atm=None):
    """utils.hpp:217-253.  q: [ncell, PCNST], written in place; atm: the
    qv, qc, qi, nc, ni arrays (None: not written)."""
    if atm is not None:
        for i in range(5):
            q[:, i] = atm[i]
    if progs["q_gas"][0] is not None:
        s = GASES_START
        for g in range(NUM_GAS):
            q[:, s] = progs["q_gas"][g]
            s += 1
    else:
        s = AERO_START
    for m in range(4):
        for a in range(NSPEC_MODE[m]):
            q[:, s] = progs["q_aero_i"][m][a]
            s += 1
        q[:, s] = progs["n_mode_i"][m]
        s += 1


def extract_ptend_from_tendencies(tends, ptend):
    """utils.hpp:255-295: the same walk without the water species."""
    extract_stateq_from_prognostics(tends, ptend, atm=None)


def inject_stateq_to_prognostics(q, progs):
    """utils.hpp:297-334.  As in the reference, the running index advances
    past a species only when its view exists (utils.hpp:318-321)."""
    if progs["q_gas"][0] is not None:
        s = GASES_START
        for g in range(NUM_GAS):
            progs["q_gas"][g] = q[:, s].copy()
            s += 1
    else:
        s = AERO_START
    for m in range(4):
        for a in range(NSPEC_MODE[m]):
            if progs["q_aero_i"][m][a] is not None:
                progs["q_aero_i"][m][a] = q[:, s].copy()
                s += 1
        progs["n_mode_i"][m] = q[:, s].copy()
        s += 1


def extract_qqcw_from_prognostics(progs, qqcw):
    """utils.hpp:364-386"""
    s = AERO_START
    for m in range(4):
        for a in range(NSPEC_MODE[m]):
            if progs["q_aero_c"][m][a] is not None:
                qqcw[:, s] = progs["q_aero_c"][m][a]
                s += 1
        qqcw[:, s] = progs["n_mode_c"][m]
        s += 1


def inject_qqcw_to_prognostics(qqcw, progs):
    """utils.hpp:388-420"""
    s = AERO_START
    for m in range(4):
        for a in range(NSPEC_MODE[m]):
            if progs["q_aero_c"][m][a] is not None:
                progs["q_aero_c"][m][a] = qqcw[:, s].copy()
                s += 1
        progs["n_mode_c"][m] = qqcw[:, s].copy()
        s += 1


def allocated_progs(ncell):
    """Prognostics with every view present (what EAMxx passes)."""
    p = empty_progs()
    z = lambda: np.zeros(ncell)
    p["q_gas"] = [z() for _ in range(NUM_GAS)]
    for m in range(4):
        p["n_mode_i"][m], p["n_mode_c"][m] = z(), z()
        for a in range(NSPEC_MODE[m]):
            p["q_aero_i"][m][a], p["q_aero_c"][m][a] = z(), z()
    return p


def progs_to_state(progs, nstate=52):
    """Prognostics -> the 52 state fields (include/mam4x.h order: numbers,
    then species mode-major in the same mode-local order, H2SO4, SOAG)."""
    ncell = len(progs["n_mode_i"][0])
    st = np.empty((nstate, ncell))
    off = [0, 7, 11, 18]
    for m in range(4):
        st[m] = progs["n_mode_i"][m]
        st[4 + m] = progs["n_mode_c"][m]
        for a in range(NSPEC_MODE[m]):
            st[8 + off[m] + a] = progs["q_aero_i"][m][a]
            st[29 + off[m] + a] = progs["q_aero_c"][m][a]
    st[50] = progs["q_gas"][2]
    st[51] = progs["q_gas"][5]
    return st


def state_to_progs(st, other_gases):
    """The 52 state fields -> Prognostics; other_gases: O3, H2O2, SO2, DMS
    (the chain does not carry them)."""
    p = allocated_progs(st.shape[1])
    off = [0, 7, 11, 18]
    for m in range(4):
        p["n_mode_i"][m] = st[m]
        p["n_mode_c"][m] = st[4 + m]
        for a in range(NSPEC_MODE[m]):
            p["q_aero_i"][m][a] = st[8 + off[m] + a]
            p["q_aero_c"][m][a] = st[29 + off[m] + a]
    o3, h2o2, so2, dms = other_gases
    p["q_gas"] = [o3, h2o2, st[50], so2, dms, st[51]]
    return p
