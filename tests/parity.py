"""Comparison helpers for the HIP chain against the CPU oracle.

Error measure: for every output plane the error of a
tendency is measured on the state it produces,

    err = |tend_gpu - tend_ref| * dt / max(|q_in|, |q_ref_post|, floor)

i.e. relative to the larger of the state before and after the step, so that a
field the step barely changes is not judged by the cancellation in its
tendency.  ``floor`` keeps values at the underflow end of the distributions
(< 1e-30 kg/kg, < 1e-10 #/kg) from dividing by ~0.  Diameters are compared
plainly relative.
"""
import numpy as np

from mam4xx_amd import layout as L


def floors():
    fl = np.empty(L.NSTATE)
    for f in range(L.NSTATE):
        fl[f] = 1e-10 if L.is_number(f) else 1e-30
    return fl


def state_rel_err(tend_x, tend_ref, state, dt):
    """Per-field max error (array [NSTATE]) and the full error array."""
    post_ref = state + dt * tend_ref
    den = np.maximum(np.maximum(np.abs(state), np.abs(post_ref)), floors()[:, None, None])
    err = np.abs(tend_x - tend_ref) * dt / den
    return err.reshape(L.NSTATE, -1).max(axis=1), err


def rel_err(x, ref, floor=0.0):
    return np.abs(x - ref) / np.maximum(np.abs(ref), floor)


def report(tend_x, tend_ref, state, dt, top=8):
    mx, err = state_rel_err(tend_x, tend_ref, state, dt)
    names = L.field_names()
    order = np.argsort(-mx)[:top]
    return ", ".join("%s=%.2e" % (names[f], mx[f]) for f in order)
