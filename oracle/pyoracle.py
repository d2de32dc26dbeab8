"""ctypes binding and build recipe for the C oracle (test infrastructure)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mam4_oracle.c")
LIB = os.path.join(HERE, "libmam4x_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


def lib_path():
    return LIB


def build(force=True):
    """Compile the oracle with gcc (-O3, OpenMP, no FMA contraction)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    cmd = ["gcc", "-O3", "-std=gnu11", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", "-Wno-unused-variable", "-o", LIB + ".tmp", SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("do_cond", ctypes.c_int),
        ("do_rename", ctypes.c_int),
        ("do_newnuc", ctypes.c_int),
        ("do_coag", ctypes.c_int),
        ("gaexch_h2so4_uptake_optaa", ctypes.c_int),
        ("newnuc_h2so4_conc_optaa", ctypes.c_int),
        ("n_so4_monolayers_pcage", ctypes.c_uint),
        ("do_calcsize", ctypes.c_int),
        ("calcsize_do_aitacc_transfer", ctypes.c_int),
        ("calcsize_do_adjust", ctypes.c_int),
        ("ntsubstep", ctypes.c_int),
    ]


def default_config(**kw):
    c = dict(do_cond=1, do_rename=1, do_newnuc=1, do_coag=1, gaexch_h2so4_uptake_optaa=2,
             newnuc_h2so4_conc_optaa=2, n_so4_monolayers_pcage=8, do_calcsize=1,
             calcsize_do_aitacc_transfer=1, calcsize_do_adjust=1, ntsubstep=1)
    c.update(kw)
    return OrcConfig(**c)


class PhotoTable(ctypes.Structure):
    """orc_photo_table: dimensions and pointers into numpy arrays that the
    instance keeps alive (photo_table())."""
    _fields_ = [(n, ctypes.c_int) for n in ("nw", "nt", "np_xs", "numj", "nump", "numsza",
                                             "numcolo3", "numalb")] + \
        [(n, ctypes.POINTER(ctypes.c_double)) for n in (
            "xsqy", "sza", "del_sza", "alb", "del_alb", "press", "del_p", "colo3", "o3rat",
            "del_o3rat", "etfphot", "prs", "dprs", "rsf_tab")] + \
        [("pht_alias_mult", ctypes.c_double), ("lng_index", ctypes.c_int)]


PHOTO_ARRAYS = ("xsqy", "sza", "del_sza", "alb", "del_alb", "press", "del_p", "colo3", "o3rat",
                "del_o3rat", "etfphot", "prs", "dprs", "rsf_tab")


def photo_table(arrays, pht_alias_mult=1.0, lng_index=0):
    """PhotoTable over a dict of numpy arrays: rsf_tab [nw, nump, numsza,
    numcolo3, numalb], xsqy [numj, nw, nt, np_xs] and the 1-D axes."""
    t = PhotoTable()
    keep = {}
    for k in PHOTO_ARRAYS:
        a = np.ascontiguousarray(arrays[k], dtype=np.float64)
        keep[k] = a
        setattr(t, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    t.numj, t.nw, t.nt, t.np_xs = keep["xsqy"].shape
    _, t.nump, t.numsza, t.numcolo3, t.numalb = keep["rsf_tab"].shape
    t.pht_alias_mult = pht_alias_mult
    t.lng_index = lng_index
    t._keep = keep
    return t


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Oracle:
    """Loads the oracle shared library (building it when missing)."""

    def __init__(self, path=None):
        path = path or LIB
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_chain_step.restype = ctypes.c_long
        L.orc_chain_step.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double] + [_dp] * 15 + [ctypes.c_int]
        L.orc_chain_step_proc.restype = ctypes.c_long
        L.orc_chain_step_proc.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double] + [_dp] * 16 + [ctypes.c_int]
        L.orc_nonvolatile_update.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_getcoags_wrapper.argtypes = [ctypes.c_double] * 10 + [_dp]
        for fn in ("orc_bm0ij", "orc_bm3i"):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [ctypes.c_int] * 3
        for fn, n in (("orc_vehk_xcrit", 3), ("orc_vehk_logj", 4), ("orc_vehk_ntot", 4),
                      ("orc_vehk_radius", 2), ("orc_vehk_threshold", 2),
                      ("orc_mean_particle_diameter_from_volume", 2),
                      ("orc_mean_particle_volume_from_diameter", 2)):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [ctypes.c_double] * n
        L.orc_nuc_mosaic_1box.argtypes = ([ctypes.c_int] + [ctypes.c_double] * 11 + [ctypes.c_int]
                                          + [ctypes.c_double] * 7 + [_dp])
        L.orc_transfer_aged_pcarbon_to_accum.argtypes = [ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_double] * 3 + [_dp] * 3
        L.orc_transfer_cond_coag_mass_to_accum.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 3
        L.orc_pcarbon_aging_1subarea.argtypes = [ctypes.c_uint] + [_dp] * 8
        L.orc_amicphys_intr.argtypes = [ctypes.POINTER(OrcConfig)] + [ctypes.c_double] * 8 + \
            [_dp] * 10 + [_dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_calcsize_level.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_double] + [_dp] * 10
        L.orc_soaexch.restype = ctypes.c_int
        L.orc_soaexch.argtypes = [ctypes.c_double] * 3 + [_dp, _dp, _dp, _dp, _dp]
        L.orc_set_libm_noise.argtypes = [ctypes.c_ulonglong]
        L.orc_set_libm_noise.restype = None
        L.orc_makoh_quartic.argtypes = [ctypes.c_double] * 4 + [_dp, _dp]
        L.orc_makoh_quartic.restype = None
        L.orc_kohler_root.argtypes = [ctypes.c_double] * 4
        L.orc_kohler_root.restype = ctypes.c_double
        L.orc_water_uptake_cell.argtypes = [_dp] + [ctypes.c_double] * 4 + [_dp] * 4
        L.orc_water_uptake_cell.restype = None
        L.orc_water_uptake.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 9
        L.orc_water_uptake.restype = None
        L.orc_qsat_water.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_qsat_water.restype = None
        L.orc_calcsize_sub_level.argtypes = [ctypes.c_double] + [_dp] * 10
        L.orc_calcsize_sub_level.restype = None
        L.orc_wet_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double] + [_dp] * 11
        L.orc_wet_size.restype = None
        L.orc_amicphys_1subarea.argtypes = ([ctypes.c_int] * 5 + [ctypes.c_double, ctypes.c_int,
                                                                  ctypes.c_uint]
                                            + [ctypes.c_double] * 5 + [_dp] * 25)
        L.orc_amicphys_1subarea.restype = ctypes.c_int
        L.orc_magnus_es.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_magnus_es.restype = ctypes.c_double
        L.orc_rename_tail_fraction.argtypes = [ctypes.c_double, ctypes.c_int, _dp]
        L.orc_rename_tail_fraction.restype = ctypes.c_double
        L.orc_set_fortran_air_constants.argtypes = [ctypes.c_int]
        L.orc_set_fortran_air_constants.restype = None

    def amicphys_1subarea(self, c, do_cond=1, do_rename=1, do_newnuc=1, do_coag=1, optaa=2,
                          nmono=8, fortran_air=False):
        """One mam_amicphys_1subarea call (mam4_amicphys.hpp:1126-1694) on the
        inputs of a dict such as the known-answer cases: gases {SOAG, H2SO4},
        masses [7][4] (amicphys species order x mode).  Returns the updated
        gases, numbers and masses, the increments with a fifth column holding
        condensation's own increments before aging, coagulation's increments
        after aging ("coag_aged"), and the SOA sub-steps.  fortran_air selects
        the Fortran MAM4 air terms of the gas diffusivity for this call."""
        f = lambda k, shape: np.ascontiguousarray(np.asarray(c[k], dtype=np.float64)  # noqa
                                                  .reshape(shape))
        z = lambda *shape: np.zeros(shape)  # noqa: E731
        dgn_a, dgn_awet, wetdens = f("dgn_a", 4), f("dgn_awet", 4), f("wetdens", 4)
        qgas1, qgas3, qnum3 = f("qgas1", 2), f("qgas3", 2), f("qnum3", 4)
        qaer2, qaer3 = f("qaer2", (7, 4)), f("qaer3", (7, 4))
        qgas, qgas_d, qnum, qnum_d = z(2), z(2, 4), z(4), z(4, 4)
        qaer, qaer_d = z(7, 4), z(7, 4, 4)
        qnumcw3, qnumcw, qnumcw_d = z(4), z(4), z(4, 1)
        qaercw2, qaercw3, qaercw, qaercw_d = z(7, 4), z(7, 4), z(7, 4), z(7, 4, 1)
        qnum_co, qaer_co, qnum_ca, qaer_ca = z(4), z(7, 4), z(4), z(7, 4)
        self.lib.orc_set_fortran_air_constants(1 if fortran_air else 0)
        try:
            n = self._subarea_call(c, optaa, do_cond, do_rename, do_newnuc, do_coag, nmono, (
                dgn_a, dgn_awet, wetdens, qgas1, qgas3, qgas, qgas_d, qnum3, qnum, qnum_d,
                qaer2, qaer3, qaer, qaer_d, qnumcw3, qnumcw, qnumcw_d, qaercw2, qaercw3,
                qaercw, qaercw_d, qnum_co, qaer_co, qnum_ca, qaer_ca))
        finally:
            self.lib.orc_set_fortran_air_constants(0)
        # gases: condensation has no aging share, so cond_only = cond
        qgas_d5 = np.concatenate([qgas_d, qgas_d[:, :1] * (do_cond != 0)], axis=1)
        qnum_d5 = np.concatenate([qnum_d, qnum_co[:, None]], axis=1)
        qaer_d5 = np.concatenate([qaer_d, qaer_co[:, :, None]], axis=2)
        return {"qgas4": qgas, "qgas_delaa": qgas_d5, "qnum4": qnum, "qnum_delaa": qnum_d5,
                "qaer4": qaer, "qaer_delaa": qaer_d5, "qnumcw4": qnumcw, "qaercw4": qaercw,
                "qnum_coag_aged": qnum_ca, "qaer_coag_aged": qaer_ca, "substeps": n}

    def _subarea_call(self, c, optaa, do_cond, do_rename, do_newnuc, do_coag, nmono, arrays):
        return self.lib.orc_amicphys_1subarea(
            optaa, do_cond, do_rename, do_newnuc, do_coag, c["deltat"], int(c["iscldy"]), nmono,
            c["temp"], c["pmid"], c["zmid"], c["pblh"], c["relhum"], *[_p(a) for a in arrays])

    # ---- aqueous sulfur chemistry (mo_setsox.hpp) ----
    def _sox_sigs(self):
        L = self.lib
        if getattr(self, "_sox_ready", False):
            return
        L.orc_sox_create_obj.argtypes = [ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double,
                                         _dp]
        L.orc_sox_create_obj.restype = None
        L.orc_sox_henry.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_sox_henry.restype = None
        L.orc_sox_ynetpos.argtypes = [ctypes.c_double, _dp] + [ctypes.c_double] * 4 + [_dp]
        L.orc_sox_ynetpos.restype = ctypes.c_double
        L.orc_sox_ph.argtypes = [ctypes.c_double] * 8 + [ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_sox_ph.restype = ctypes.c_double
        L.orc_sox_aer_factor.argtypes = [_dp, _dp]
        L.orc_sox_aer_factor.restype = None
        L.orc_sox_uptakerate.argtypes = [ctypes.c_double] * 6
        L.orc_sox_uptakerate.restype = ctypes.c_double
        L.orc_setsox_single_level.argtypes = [ctypes.c_double] * 7 + [_dp] * 5
        L.orc_setsox_single_level.restype = None
        L.orc_setsox.argtypes = [ctypes.c_long, ctypes.c_double] + [_dp] * 10
        L.orc_setsox.restype = None
        self._sox_ready = True

    def sox_create_obj(self, cldfrc, qcw, lwc, cfact):
        self._sox_sigs()
        out = np.zeros(3)
        q = np.ascontiguousarray(qcw, dtype=np.float64)
        self.lib.orc_sox_create_obj(cldfrc, _p(q), lwc, cfact, _p(out))
        return out

    def sox_henry(self, species, t_factor):
        """0 SO2 (xk, xe, x2), 1 CO2 (xk, xe), 2 H2O2 (xk, xe), 3 O3 (xk)"""
        self._sox_sigs()
        out = np.zeros(3)
        self.lib.orc_sox_henry(species, t_factor, _p(out))
        return out

    def sox_ynetpos(self, yph, f, Eco2, Eh2o, Eso4, so4_fact=1.0):
        self._sox_sigs()
        xph = np.zeros(1)
        fa = np.ascontiguousarray(f, dtype=np.float64)
        y = self.lib.orc_sox_ynetpos(yph, _p(fa), Eco2, Eh2o, Eso4, so4_fact, _p(xph))
        return y, xph[0]

    def sox_ph(self, T, patm, xlwc, t_factor, xso2, xso4, xhnm, so4_fact=1.0):
        """([H+], converged, factors[7] = fact1-4_so2, Eco2, Eh2o, Eso4)"""
        self._sox_sigs()
        conv = ctypes.c_int(0)
        f = np.zeros(7)
        x = self.lib.orc_sox_ph(T, patm, xlwc, t_factor, xso2, xso4, xhnm, so4_fact,
                                ctypes.byref(conv), _p(f))
        return x, bool(conv.value), f

    def sox_aer_factor(self, tmr):
        self._sox_sigs()
        t = np.ascontiguousarray(tmr, dtype=np.float64)
        out = np.zeros(4)
        self.lib.orc_sox_aer_factor(_p(t), _p(out))
        return out

    def sox_uptakerate(self, xl, cldnum, cfact, cldfrc, tfld, press):
        self._sox_sigs()
        return self.lib.orc_sox_uptakerate(xl, cldnum, cfact, cldfrc, tfld, press)

    def setsox_single_level(self, dt, press, tfld, lwc, cldfrc, cldnum, xhnm, qcw, qin):
        """one cell: returns (qcw, qin, dqdt_aqso4, dqdt_aqh2so4, [H+])"""
        self._sox_sigs()
        qc = np.array(qcw, dtype=np.float64)
        qi = np.array(qin, dtype=np.float64)
        d1, d2, ph = np.zeros(31), np.zeros(31), np.zeros(1)
        self.lib.orc_setsox_single_level(dt, press, tfld, lwc, cldfrc, cldnum, xhnm, _p(qc),
                                         _p(qi), _p(d1), _p(d2), _p(ph))
        return qc, qi, d1, d2, ph[0]

    def setsox(self, dt, met, qcw, qin):
        """a domain: met dict of planes (press, tfld, lwc, cldfrc, cldnum,
        xhnm; any common shape), qcw/qin [31, ...]; returns (qcw, qin,
        dqdt_aqso4, dqdt_aqh2so4)"""
        self._sox_sigs()
        shape = np.shape(met["tfld"])
        n = int(np.prod(shape))
        m = [np.ascontiguousarray(met[k], dtype=np.float64).reshape(n)
             for k in ("press", "tfld", "lwc", "cldfrc", "cldnum", "xhnm")]
        qc = np.array(qcw, dtype=np.float64).reshape(31, n)
        qi = np.array(qin, dtype=np.float64).reshape(31, n)
        d1, d2 = np.zeros((31, n)), np.zeros((31, n))
        self.lib.orc_setsox(n, dt, *[_p(a) for a in m], _p(qc), _p(qi), _p(d1), _p(d2))
        sh = (31,) + tuple(shape)
        return qc.reshape(sh), qi.reshape(sh), d1.reshape(sh), d2.reshape(sh)

    # ---- gas-phase chemistry (gas_chem.hpp, gas_phase_chemistry.hpp) ----
    def _gc_sigs(self):
        L = self.lib
        if getattr(self, "_gc_ready", False):
            return
        L.orc_gc_rates.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_gc_rates.restype = None
        L.orc_gc_imp_sol.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_gc_imp_sol.restype = ctypes.c_int
        L.orc_gas_phase_chemistry.argtypes = [ctypes.c_double] * 3 + [_dp] * 4
        L.orc_gas_phase_chemistry.restype = ctypes.c_int
        L.orc_gas_phase_chemistry_domain.argtypes = [ctypes.c_long, ctypes.c_double] + [_dp] * 6 + \
            [ctypes.POINTER(ctypes.c_int)]
        L.orc_gas_phase_chemistry_domain.restype = None
        self._gc_ready = True

    def gc_rates(self, temp, inv):
        """reaction rates rxt[7] after setrxt/usrrxt/adjrxt (rxt[0], the
        photolysis rate, left 0)"""
        self._gc_sigs()
        out = np.zeros(7)
        v = np.ascontiguousarray(inv, dtype=np.float64)
        self.lib.orc_gc_rates(temp, _p(v), _p(out))
        return out

    def gc_imp_sol(self, base_sol, rxt, het, extfrc, delt):
        """imp_sol on one cell: (new base_sol[31], time steps)"""
        self._gc_sigs()
        y = np.array(base_sol, dtype=np.float64)
        r, h, e = (np.ascontiguousarray(a, dtype=np.float64) for a in (rxt, het, extfrc))
        n = self.lib.orc_gc_imp_sol(_p(y), _p(r), _p(h), _p(e), delt)
        return y, n

    def gas_phase_chemistry(self, dt, temp, photo, extfrc, inv, het, qq):
        """a domain: temp/photo [...], extfrc [9, ...], inv [8, ...], het /
        qq [31, ...]; returns (qq, steps)"""
        self._gc_sigs()
        shape = np.shape(temp)
        n = int(np.prod(shape))
        t = np.ascontiguousarray(temp, dtype=np.float64).reshape(n)
        ph = np.ascontiguousarray(photo, dtype=np.float64).reshape(n)
        e = np.ascontiguousarray(extfrc, dtype=np.float64).reshape(9, n)
        v = np.ascontiguousarray(inv, dtype=np.float64).reshape(8, n)
        h = np.ascontiguousarray(het, dtype=np.float64).reshape(31, n)
        q = np.array(qq, dtype=np.float64).reshape(31, n)
        st = np.zeros(n, dtype=np.int32)
        self.lib.orc_gas_phase_chemistry_domain(n, dt, _p(t), _p(ph), _p(e), _p(v), _p(h), _p(q),
                                                st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return q.reshape((31,) + tuple(shape)), st.reshape(shape)

    # ---- photolysis (mo_photo.hpp) ----
    def _pht_sigs(self):
        L = self.lib
        if getattr(self, "_pht_ready", False):
            return
        D, I, T = ctypes.c_double, ctypes.c_int, ctypes.POINTER(PhotoTable)
        L.orc_photo_find_index.argtypes = [_dp, I, D, I]
        L.orc_photo_find_index.restype = I
        L.orc_photo_cloud_mod.argtypes = [I, D, _dp, _dp, _dp, D, _dp, _dp]
        L.orc_photo_cloud_mod.restype = None
        L.orc_photo_interpolate_rsf.argtypes = [T, I, _dp, D, _dp, _dp, _dp]
        L.orc_photo_interpolate_rsf.restype = None
        L.orc_photo_jlong.argtypes = [T, I, D, _dp, _dp, _dp, _dp, _dp]
        L.orc_photo_jlong.restype = None
        L.orc_table_photo.argtypes = [T, I, _dp, _dp, _dp, _dp, _dp, D, D, _dp, _dp, D]
        L.orc_table_photo.restype = None
        L.orc_table_photo_domain.argtypes = [T, I, I] + [_dp] * 9 + [D]
        L.orc_table_photo_domain.restype = None
        self._pht_ready = True

    def photo_find_index(self, var, vmin, idx0=0):
        self._pht_sigs()
        v = np.ascontiguousarray(var, dtype=np.float64)
        return self.lib.orc_photo_find_index(_p(v), len(v), vmin, idx0)

    def photo_cloud_mod(self, zen_angle, clouds, lwc, delp, srf_alb):
        """one column [nlev]: (eff_alb, cld_mult)"""
        self._pht_sigs()
        c, w, d = (np.ascontiguousarray(a, dtype=np.float64) for a in (clouds, lwc, delp))
        ea, cm = np.empty_like(c), np.empty_like(c)
        self.lib.orc_photo_cloud_mod(len(c), zen_angle, _p(c), _p(w), _p(d), srf_alb, _p(ea),
                                     _p(cm))
        return ea, cm

    def photo_interpolate_rsf(self, table, alb_in, sza_in, p_in, colo3_in):
        """one column: rsf [nw, nlev]; table: a PhotoTable (photo_table())"""
        self._pht_sigs()
        a, p, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (alb_in, p_in, colo3_in))
        out = np.empty((table.nw, len(a)))
        self.lib.orc_photo_interpolate_rsf(ctypes.byref(table), len(a), _p(a), sza_in, _p(p),
                                           _p(c), _p(out))
        return out

    def photo_jlong(self, table, sza_in, alb_in, p_in, t_in, colo3_in):
        """one column: j_long [numj, nlev]"""
        self._pht_sigs()
        a, p, t, c = (np.ascontiguousarray(x, dtype=np.float64)
                      for x in (alb_in, p_in, t_in, colo3_in))
        out = np.empty((table.numj, len(a)))
        self.lib.orc_photo_jlong(ctypes.byref(table), len(a), sza_in, _p(a), _p(p), _p(t),
                                 _p(c), _p(out))
        return out

    def table_photo(self, table, cols, esfact=1.0):
        """columns: dict of photo, pmid, pdel, temp, colo3, lwc, clouds
        [nlev, ncol] and zen_angle, srf_alb [ncol].  Returns the new photo
        [nlev, ncol]."""
        self._pht_sigs()
        nlev, ncol = np.shape(cols["photo"])
        f = lambda k: np.ascontiguousarray(cols[k], dtype=np.float64)
        ph = np.array(cols["photo"], dtype=np.float64)
        self.lib.orc_table_photo_domain(ctypes.byref(table), ncol, nlev, _p(ph),
                                        *[_p(f(k)) for k in ("pmid", "pdel", "temp", "colo3",
                                                             "zen_angle", "srf_alb", "lwc",
                                                             "clouds")], esfact)
        return ph

    # ---- stratospheric ozone and invariants (lin_strat_chem.hpp, mo_setinv.hpp) ----
    def _lnz_sigs(self):
        L = self.lib
        if getattr(self, "_lnz_ready", False):
            return
        D, I = ctypes.c_double, ctypes.c_int
        L.orc_o3_column_density.argtypes = [I, _dp, _dp, D, _dp]
        L.orc_o3_column_density.restype = None
        L.orc_linoz_solve_kk.argtypes = [D] * 6 + [_dp, D, D, _dp, _dp]
        L.orc_linoz_solve_kk.restype = None
        L.orc_linoz_sfcsink.argtypes = [I, D, _dp, _dp, D, I, D]
        L.orc_linoz_sfcsink.restype = D
        L.orc_linoz_column.argtypes = [I, D, D, D, I, D, D, D, D, D, I] + [_dp] * 10
        L.orc_linoz_column.restype = None
        L.orc_linoz_domain.argtypes = [I, I, D, D, D, D, D, I] + [_dp] * 7 + \
            [ctypes.POINTER(ctypes.c_int)] + [_dp] * 6
        L.orc_linoz_domain.restype = None
        L.orc_setinv_single_level.argtypes = [D, D, D, _dp, _dp]
        L.orc_setinv_single_level.restype = None
        L.orc_setinv.argtypes = [ctypes.c_long] + [_dp] * 5
        L.orc_setinv.restype = None
        self._lnz_ready = True

    def o3_column_density(self, pdel, vmr_o3, col0=0.0):
        self._lnz_sigs()
        p, v = (np.ascontiguousarray(a, dtype=np.float64) for a in (pdel, vmr_o3))
        out = np.empty_like(p)
        self.lib.orc_o3_column_density(len(p), _p(p), _p(v), col0, _p(out))
        return out

    def linoz_solve_kk(self, o3col, temp, sza, pmid, dt, rlats, clim, chlorine_loading, psc_T,
                       o3):
        """one level: (new o3, [do3_linoz, do3_linoz_psc, ss_o3])"""
        self._lnz_sigs()
        c = np.ascontiguousarray(clim, dtype=np.float64)
        o = np.array([o3], dtype=np.float64)
        d = np.zeros(3)
        self.lib.orc_linoz_solve_kk(o3col, temp, sza, pmid, dt, rlats, _p(c), chlorine_loading,
                                    psc_T, _p(o), _p(d))
        return o[0], d

    def linoz_sfcsink(self, dt, pdel, o3, o3_sfc, o3_lbl, o3_tau):
        """(new o3 column, sink [Tg/yr/m2])"""
        self._lnz_sigs()
        p = np.ascontiguousarray(pdel, dtype=np.float64)
        o = np.array(o3, dtype=np.float64)
        s = self.lib.orc_linoz_sfcsink(len(p), dt, _p(p), _p(o), o3_sfc, o3_lbl, o3_tau)
        return o, s

    def linoz(self, dt, cols, conf):
        """columns: dict of temp, pmid, pdel, o3 [nlev, ncol], clim [8, nlev,
        ncol], sza, rlats, col0 [ncol], ltrop (int) [ncol]; conf: dict of
        chlorine_loading, psc_T, o3_sfc, o3_tau, o3_lbl.  Returns (o3, o3col,
        do3, do3_psc, ss_o3 [nlev, ncol], sfcsink [ncol])."""
        self._lnz_sigs()
        nlev, ncol = np.shape(cols["o3"])
        f = lambda k: np.ascontiguousarray(cols[k], dtype=np.float64)
        o3 = np.array(cols["o3"], dtype=np.float64)
        outs = [np.empty((nlev, ncol)) for _ in range(4)]
        sink = np.empty(ncol)
        lt = np.ascontiguousarray(cols["ltrop"], dtype=np.int32)
        self.lib.orc_linoz_domain(ncol, nlev, dt, conf["chlorine_loading"], conf["psc_T"],
                                  conf["o3_sfc"], conf["o3_tau"], int(conf["o3_lbl"]),
                                  *[_p(f(k)) for k in ("temp", "pmid", "pdel", "clim", "sza",
                                                       "rlats", "col0")],
                                  lt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(o3),
                                  *[_p(b) for b in outs], _p(sink))
        return (o3, *outs, sink)

    def setinv(self, tfld, pmid, qv, cnst_offline):
        """invariants [8, ...] [#/cm3] from T, p, qv (kg/kg) and the four
        offline species [4, ...] (vmr)"""
        self._lnz_sigs()
        shape = np.shape(tfld)
        n = int(np.prod(shape))
        t, p, q = (np.ascontiguousarray(a, dtype=np.float64).reshape(n) for a in (tfld, pmid, qv))
        co = np.ascontiguousarray(cnst_offline, dtype=np.float64).reshape(4, n)
        inv = np.empty((8, n))
        self.lib.orc_setinv(n, _p(t), _p(p), _p(q), _p(co), _p(inv))
        return inv.reshape((8,) + tuple(shape))

    # ---- washout rates (mo_sethet.hpp) ----
    def _sh_sigs(self):
        L = self.lib
        if getattr(self, "_sh_ready", False):
            return
        D, I = ctypes.c_double, ctypes.c_int
        L.orc_precip_rescale.argtypes = [I, _dp, _dp, _dp, _dp]
        L.orc_precip_rescale.restype = None
        L.orc_gas_washout.argtypes = [I, I, D] + [_dp] * 6
        L.orc_gas_washout.restype = None
        L.orc_find_ktop.argtypes = [I, D, _dp]
        L.orc_find_ktop.restype = I
        L.orc_sethet_xkgm.argtypes = []
        L.orc_sethet_xkgm.restype = D
        L.orc_sethet_column.argtypes = [I, D, D, D] + [_dp] * 10
        L.orc_sethet_column.restype = None
        L.orc_sethet_domain.argtypes = [I, I, D] + [_dp] * 12
        L.orc_sethet_domain.restype = None
        self._sh_ready = True

    def precip_rescale(self, cmfdqr, nrain, nevapr):
        self._sh_sigs()
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cmfdqr, nrain, nevapr)]
        out = np.empty_like(a[0])
        self.lib.orc_precip_rescale(len(out), *[_p(x) for x in a], _p(out))
        return out

    def gas_washout(self, ktop, xkgm, xliq, rain, xhen, tfld, delz, xgas):
        self._sh_sigs()
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (xliq, rain, xhen, tfld, delz)]
        g = np.array(xgas, dtype=np.float64)
        self.lib.orc_gas_washout(len(g), ktop, xkgm, *[_p(x) for x in a], _p(g))
        return g

    def find_ktop(self, rlat, press):
        self._sh_sigs()
        p = np.ascontiguousarray(press, dtype=np.float64)
        return self.lib.orc_find_ktop(len(p), rlat, _p(p))

    def sethet_xkgm(self):
        self._sh_sigs()
        return self.lib.orc_sethet_xkgm()

    def sethet(self, delt, cols):
        """cols: dict of press, zmid, tfld, cmfdqr, nrain, nevapr, inv_m,
        h2o2, so2 [nlev, ncol] and rlat, phis [ncol]; returns het_rates
        [31, nlev, ncol] [1/s]"""
        self._sh_sigs()
        nlev, ncol = np.shape(cols["press"])
        f = lambda k: np.ascontiguousarray(cols[k], dtype=np.float64)
        het = np.empty((31, nlev, ncol))
        self.lib.orc_sethet_domain(ncol, nlev, delt, _p(f("rlat")), _p(f("phis")),
                                   *[_p(f(k)) for k in ("press", "zmid", "tfld", "cmfdqr",
                                                        "nrain", "nevapr", "inv_m", "h2o2",
                                                        "so2")], _p(het))
        return het

    # ---- external forcing (mo_setext.hpp) ----
    def setext(self, sectors, frc_ndx=None, file_alt_data=None):
        """sectors: 9 lists (one per forced species) of up to 4 [nlev, ncol]
        arrays; returns extfrc [9, nlev, ncol] (entries no forcing targets
        are 0)"""
        L = self.lib
        L.orc_setext_column.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER((ctypes.c_void_p * 4) * 9), _dp]
        L.orc_setext_column.restype = None
        shape = next(np.shape(s[0]) for s in sectors if len(s))
        nlev, ncol = shape
        ndx = (ctypes.c_int * 9)(*(frc_ndx if frc_ndx is not None else range(1, 10)))
        alt = (ctypes.c_int * 9)(*([int(bool(x)) for x in file_alt_data]
                                   if file_alt_data is not None else [0] * 9))
        nsec = (ctypes.c_int * 9)(*[len(s) for s in sectors])
        out = np.zeros((9, nlev, ncol))
        for j in range(ncol):
            cols = [[np.ascontiguousarray(np.asarray(x, dtype=np.float64)[:, j]) for x in s]
                    for s in sectors]
            tab = ((ctypes.c_void_p * 4) * 9)()
            for m in range(9):
                for i, c in enumerate(cols[m]):
                    tab[m][i] = c.ctypes.data
            fr = np.zeros((9, nlev))
            L.orc_setext_column(nlev, ndx, alt, nsec, ctypes.byref(tab), _p(fr))
            out[:, :, j] = fr
        return out

    def set_libm_noise(self, seed):
        """seed != 0: randomly move each exp/log/pow/cbrt/erfc result by one
        ulp (conditioning probe); 0: exact libm."""
        self.lib.orc_set_libm_noise(int(seed))

    def envelope(self, scene, dt=1800.0, config=None, seeds=(1, 2, 3, 4, 5, 6), return_dgn=False,
                 **kw):
        """Per-element spread of the tendencies under libm-rounding noise,
        in the state-relative measure of tests/parity.py.  Returns
        (base tendencies, envelope array [NSTATE, nlev, ncol]) and, with
        return_dgn, the relative spread of the calcsize diameters [2, 4, nlev, ncol]."""
        import numpy as np
        base, bdi, bdc, _ = self.chain_step(scene, dt=dt, config=config, **kw)
        st = scene["state"]
        post = st + dt * base
        fl = np.where(np.arange(st.shape[0]) < 8, 1e-10, 1e-30)[:, None, None]
        den = np.maximum(np.maximum(np.abs(st), np.abs(post)), fl)
        env = np.zeros_like(base)
        env_dgn = np.zeros((2,) + bdi.shape)
        try:
            for s in seeds:
                self.set_libm_noise(s)
                t, di, dc, _ = self.chain_step(scene, dt=dt, config=config, **kw)
                env = np.maximum(env, np.abs(t - base) * dt / den)
                env_dgn[0] = np.maximum(env_dgn[0], np.abs(di - bdi) / np.abs(bdi))
                env_dgn[1] = np.maximum(env_dgn[1], np.abs(dc - bdc) / np.abs(bdc))
        finally:
            self.set_libm_noise(0)
        if return_dgn:
            return base, env, env_dgn
        return base, env

    def chain_step(self, scene, dt=1800.0, config=None, nthreads=0, pregaschem_gas=None,
                   precldchem=None, proc=False):
        """Run the composed step on a scene dict; returns (tend, dgn_i, dgn_c, soa_iters)
        and, with proc, the per-process tendency diagnostics [5, 31, nlev, ncol]
        (condensation, renaming, nucleation, coagulation, cloud-borne renaming;
        gas_pcnst tracers, mol/mol/s or #/kmol/s) as a fifth element."""
        cfg = config or default_config()
        st = np.ascontiguousarray(scene["state"], dtype=np.float64)
        nf, nlev, ncol = st.shape
        tend = np.empty_like(st)
        dgn_i = np.empty((4, nlev, ncol))
        dgn_c = np.empty((4, nlev, ncol))
        arrs = [np.ascontiguousarray(scene[k], dtype=np.float64) for k in
                ("temperature", "pressure", "height", "qv", "cloud_fraction", "pblh",
                 "dgncur_a", "dgncur_awet", "wetdens")]
        pg = None if pregaschem_gas is None else np.ascontiguousarray(pregaschem_gas)
        pc = None if precldchem is None else np.ascontiguousarray(precldchem)
        pr = np.empty((5, 31, nlev, ncol)) if proc else None
        iters = self.lib.orc_chain_step_proc(ctypes.byref(cfg), ncol, nlev, dt, _p(st),
                                             *[_p(a) for a in arrs], _p(pg), _p(pc), _p(tend),
                                             _p(dgn_i), _p(dgn_c), _p(pr), int(nthreads))
        if proc:
            return tend, dgn_i, dgn_c, int(iters), pr
        return tend, dgn_i, dgn_c, int(iters)

    # ---- process-level helpers for the known-answer tests ----
    def makoh_quartic(self, p3, p2, p1, p0):
        """The four complex roots of x^4 + p3 x^3 + p2 x^2 + p1 x + p0."""
        re, im = np.zeros(4), np.zeros(4)
        self.lib.orc_makoh_quartic(p3, p2, p1, p0, _p(re), _p(im))
        return re + 1j * im

    def kohler_root(self, aa, rdry, hygro, rh):
        return self.lib.orc_kohler_root(aa, rdry, hygro, rh)

    def qsat_water(self, T, p):
        """(es [Pa], qs) of Goff & Gratch over water."""
        es, qs = np.zeros(1), np.zeros(1)
        self.lib.orc_qsat_water(T, p, _p(es), _p(qs))
        return es[0], qs[0]

    def water_uptake_cell(self, qaer, T, p, qv, cld, dgncur_a):
        """One cell: qaer [4][7] kg/kg -> (dgncur_awet[4], qaerwat[4], wetdens[4])."""
        qa = np.ascontiguousarray(qaer, dtype=np.float64).reshape(4, 7)
        dga = np.ascontiguousarray(dgncur_a, dtype=np.float64)
        out = [np.zeros(4) for _ in range(3)]
        self.lib.orc_water_uptake_cell(_p(qa), T, p, qv, cld, _p(dga), *[_p(o) for o in out])
        return tuple(out)

    def water_uptake(self, scene, dgncur_a=None):
        """Water uptake over a scene dict ([field][level][column] arrays);
        dgncur_a defaults to the scene's.  Returns (dgncur_awet, qaerwat,
        wetdens), each [4, nlev, ncol]."""
        st = np.ascontiguousarray(scene["state"])
        _, nlev, ncol = st.shape
        dga = np.ascontiguousarray(scene["dgncur_a"] if dgncur_a is None else dgncur_a)
        out = [np.empty((4, nlev, ncol)) for _ in range(3)]
        c = lambda k: np.ascontiguousarray(scene[k])
        self.lib.orc_water_uptake(ncol, nlev, _p(st), _p(c("temperature")), _p(c("pressure")),
                                  _p(c("qv")), _p(c("cloud_fraction")), _p(dga),
                                  *[_p(o) for o in out])
        return tuple(out)

    def calcsize_level(self, dt, n_i, n_c, q_i, q_c, e3sm=False, config=None):
        """One level of calcsize: CalcSize::compute_tendencies (e3sm=False)
        or modal_aero_calcsize_sub (e3sm=True).  Returns (dn_i, dn_c, dq_i,
        dq_c, dgn_i, dgn_c)."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (n_i, n_c)]
        q = [np.ascontiguousarray(x, dtype=np.float64).reshape(4, 7) for x in (q_i, q_c)]
        dn_i, dn_c = np.zeros(4), np.zeros(4)
        dq_i, dq_c = np.zeros((4, 7)), np.zeros((4, 7))
        di, dc = np.zeros(4), np.zeros(4)
        outs = [_p(x) for x in (dn_i, dn_c, dq_i, dq_c, di, dc)]
        if e3sm:
            self.lib.orc_calcsize_sub_level(dt, _p(a[0]), _p(a[1]), _p(q[0]), _p(q[1]), *outs)
        else:
            cfg = config or default_config()
            self.lib.orc_calcsize_level(ctypes.byref(cfg), dt, _p(a[0]), _p(a[1]), _p(q[0]),
                                        _p(q[1]), *outs)
        return dn_i, dn_c, dq_i, dq_c, di, dc

    def wet_size(self, scene, dt):
        """The wet-size producer over a scene: (dgn_i, dgn_c, dgncur_awet,
        qaerwat, wetdens, tend)."""
        st = np.ascontiguousarray(scene["state"])
        nf, nlev, ncol = st.shape
        outs = [np.empty((4, nlev, ncol)) for _ in range(5)] + [np.empty_like(st)]
        c = lambda k: np.ascontiguousarray(scene[k])
        self.lib.orc_wet_size(ncol, nlev, dt, _p(st), _p(c("temperature")), _p(c("pressure")),
                              _p(c("qv")), _p(c("cloud_fraction")), *[_p(o) for o in outs])
        return tuple(outs)

    def soaexch(self, dt, temp, pmid, q_soag, soa, pom, uptk_soag):
        """SOA exchange of one subarea; returns (q_soag, q_soag_avg, soa[4], niter)."""
        g = np.array([q_soag], dtype=np.float64)
        gavg = np.zeros(1)
        a = np.array(soa, dtype=np.float64)
        po = np.ascontiguousarray(pom, dtype=np.float64)
        up = np.ascontiguousarray(uptk_soag, dtype=np.float64)
        it = self.lib.orc_soaexch(dt, temp, pmid, _p(g), _p(gavg), _p(a), _p(po), _p(up))
        return g[0], gavg[0], a, it

    def nonvolatile_update(self, dt, netprod, uptkaer, qgas_cur, qaer_cur):
        up = np.ascontiguousarray(uptkaer, dtype=np.float64)
        qa = np.array(qaer_cur, dtype=np.float64)
        qg = ctypes.c_double(qgas_cur)
        qavg = ctypes.c_double(0.0)
        self.lib.orc_nonvolatile_update(dt, netprod, _p(up), ctypes.byref(qg), ctypes.byref(qavg),
                                        _p(qa))
        return qg.value, qavg.value, qa

    def getcoags(self, T, p, dgatk, dgacc, sgatk, sgacc, rhoat, rhoac):
        import math
        out = np.zeros(4)
        self.lib.orc_getcoags_wrapper(T, p, dgatk, dgacc, sgatk, sgacc, math.log(sgatk),
                                      math.log(sgacc), rhoat, rhoac, _p(out))
        return out
