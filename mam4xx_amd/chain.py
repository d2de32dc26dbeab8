"""Python host interface over the C ABI (include/mam4x.h).

``Chain`` owns one library handle; ``compute_tendencies`` takes torch tensors
resident on the GPU in the ``[field][level][column]`` layout and enqueues the
step on torch's current stream.  There is no CPU path: the HIP library is
required, and a missing library is an error.
"""
import ctypes

import torch

from . import _abi
from . import layout as L


class ChainError(RuntimeError):
    pass


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _plane_ptrs(t, n):
    """n plane pointers of a [n][nlev][ncol] tensor."""
    assert t.dim() == 3 and t.shape[0] == n, t.shape
    assert t.dtype == torch.float64 and t.is_cuda
    assert t.stride(2) == 1 and t.stride(1) == t.shape[2]
    base = t.data_ptr()
    step = t.stride(0) * 8
    return [base + i * step for i in range(n)]


def _process_ptrs(process):
    """Block pointers and tracer-plane stride of the per-process diagnostics:
    a [5][31][nlev][ncol] tensor, or a list of five [31][nlev][ncol] tensors
    (None = block not formed) sharing one tracer-plane stride."""
    blocks = list(process) if isinstance(process, (list, tuple)) else \
        [process[k] for k in range(_abi.NPROCESS)]
    assert len(blocks) == _abi.NPROCESS
    ptrs, stride = [], None
    for b in blocks:
        if b is None:
            ptrs.append(None)
            continue
        assert b.dim() == 3 and b.shape[0] == _abi.GAS_PCNST, b.shape
        assert b.dtype == torch.float64 and b.is_cuda
        assert b.stride(2) == 1 and b.stride(1) == b.shape[2]
        assert stride in (None, b.stride(0)), "blocks must share one tracer-plane stride"
        stride = b.stride(0)
        ptrs.append(b.data_ptr())
    return ptrs, stride or 0


def default_config(**kw):
    cfg = _abi.Config()
    _abi.load().mam4x_default_config(ctypes.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class Chain:
    """MAM4 microphysics chain on one GPU (AeroProcess-style interface)."""

    def __init__(self, config=None, precision=_abi.PREC_FP64, **kw):
        self.lib = _abi.load()
        cfg = config or default_config(**kw)
        cfg.precision = precision
        h = ctypes.c_void_p()
        rc = self.lib.mam4x_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != _abi.OK:
            raise ChainError("mam4x_create failed (%d)" % rc)
        self.h = h
        self.config = cfg

    def init(self, config=None, **kw):
        """Re-initialize the handle with a new configuration (mam4x_init,
        AeroProcess::init): an _abi.Config, or the defaults updated by kw."""
        cfg = config or default_config(**kw)
        self._check(self.lib.mam4x_init(self.h, ctypes.byref(cfg)), "mam4x_init")
        self.config = cfg

    def close(self):
        if getattr(self, "h", None):
            self.lib.mam4x_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != _abi.OK:
            msg = self.lib.mam4x_last_error(self.h)
            raise ChainError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    @staticmethod
    def make_inputs(scene, precldchem=None, pregaschem=None, layout=_abi.LAYOUT_LEV_COL,
                    live=None):
        """Build an _abi.Inputs from a dict of CUDA float64 tensors (see
        scenes.generate for the keys): planes [field][nlev][ncol] for
        LAYOUT_LEV_COL, [field][ncol][nlev] for LAYOUT_COL_LEV.  The planes'
        fastest dimension is the leading dimension; live (default: all of
        it) is how many of its entries the domain uses."""
        st = scene["state"]
        nf = st.shape[0]
        assert nf == L.NSTATE
        inp = _abi.Inputs()
        fast = st.shape[2] if live is None else int(live)
        assert 0 <= fast <= st.shape[2], (fast, st.shape)
        if layout == _abi.LAYOUT_LEV_COL:
            nlev, ncol = st.shape[1], fast
        else:
            ncol, nlev = st.shape[1], fast
        inp.ncol, inp.nlev, inp.ldc, inp.layout = ncol, nlev, st.shape[2], layout
        for i, p in enumerate(_plane_ptrs(st, L.NSTATE)):
            inp.state[i] = p
        # a missing key leaves its pointer NULL (the library reports what a
        # call needs)
        for key, fld in (("temperature", "temperature"), ("pressure", "pressure"),
                         ("height", "height"), ("qv", "vapor_mixing_ratio"),
                         ("cloud_fraction", "cloud_fraction"), ("pblh", "pblh")):
            t = scene.get(key)
            if t is None:
                continue
            assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous()
            setattr(inp, fld, t.data_ptr())
        for key in ("dgncur_a", "dgncur_awet", "wetdens"):
            if scene.get(key) is None:
                continue
            arr = getattr(inp, key)
            for m, p in enumerate(_plane_ptrs(scene[key], 4)):
                arr[m] = p
        if precldchem is not None:
            for i, p in enumerate(_plane_ptrs(precldchem, L.NSTATE)):
                inp.precldchem[i] = p
        if pregaschem is not None:
            pg = _plane_ptrs(pregaschem, 2)
            inp.pregaschem_h2so4, inp.pregaschem_soag = pg[0], pg[1]
        dp = scene.get("hydrostatic_dp")
        if dp is not None:
            assert dp.dtype == torch.float64 and dp.is_cuda and dp.is_contiguous()
            inp.hydrostatic_dp = dp.data_ptr()
        return inp

    def compute_tendencies(self, scene, dt, tend, dgn_i=None, dgn_c=None, state_out=None,
                           precldchem=None, pregaschem=None, t=0.0, stream=None,
                           layout=_abi.LAYOUT_LEV_COL, process=None, noalloc=False,
                           soa_capped=None, budgets=None, live=None):
        """One step (mam4x_compute_tendencies, or with noalloc the
        allocation-free mam4x_compute_tendencies_noalloc).  process: optional
        per-process tendency diagnostics, [5][31][nlev][ncol] (see
        _process_ptrs).  soa_capped: optional int32 [nlev][ncol], 1 where a
        SOA loop stopped at its 1000-sub-step cap.  budgets: optional float64
        [NBUDGET][ncol], the column budgets of the step's tendencies (needs
        scene["hydrostatic_dp"]).  live: see make_inputs."""
        inp = self.make_inputs(scene, precldchem, pregaschem, layout, live)
        out = _abi.Outputs()
        if budgets is not None:
            assert budgets.dtype == torch.float64 and budgets.is_cuda
            assert budgets.is_contiguous() and budgets.shape == (_abi.NBUDGET, inp.ncol)
            out.column_budgets = budgets.data_ptr()
        if soa_capped is not None:
            assert soa_capped.dtype == torch.int32 and soa_capped.is_cuda
            assert soa_capped.is_contiguous()
            out.soa_capped = soa_capped.data_ptr()
        if process is not None:
            ptrs, out.process_stride = _process_ptrs(process)
            for k, p in enumerate(ptrs):
                out.process_tendencies[k] = p
        for i, p in enumerate(_plane_ptrs(tend, L.NSTATE)):
            out.tendencies[i] = p
        if dgn_i is not None:
            for m, p in enumerate(_plane_ptrs(dgn_i, 4)):
                out.dgncur_i[m] = p
        if dgn_c is not None:
            for m, p in enumerate(_plane_ptrs(dgn_c, 4)):
                out.dgncur_c[m] = p
        if state_out is not None:
            for i, p in enumerate(_plane_ptrs(state_out, L.NSTATE)):
                out.state_out[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        name = "mam4x_compute_tendencies_noalloc" if noalloc else "mam4x_compute_tendencies"
        rc = getattr(self.lib, name)(self.h, t, dt, ctypes.byref(inp), ctypes.byref(out),
                                     ctypes.c_void_p(s.cuda_stream))
        self._check(rc, name)

    @staticmethod
    def _tracers(state_q, qqcw):
        """_abi.Tracers of (ncol, nlev, PCNST) tensors, tracer index fastest
        (any column / level strides, e.g. a permuted level-major array)"""
        for t in (state_q, qqcw):
            if t is None:
                continue
            assert t.dim() == 3 and t.shape[2] == _abi.PCNST, t.shape
            assert t.dtype == torch.float64 and t.is_cuda and t.stride(2) == 1
        if qqcw is not None:
            assert qqcw.shape == state_q.shape and qqcw.stride() == state_q.stride()
        tr = _abi.Tracers()
        tr.state_q = state_q.data_ptr()
        tr.qqcw = qqcw.data_ptr() if qqcw is not None else None
        tr.col_stride, tr.lev_stride = state_q.stride(0), state_q.stride(1)
        return tr

    @staticmethod
    def _fields(fields, ncol, nlev, layout):
        """_abi.Fields of a [52][nlev][ncol] ([52][ncol][nlev] for
        LAYOUT_COL_LEV) tensor, or a list of 52 planes / None"""
        fl = _abi.Fields()
        fl.ncol, fl.nlev, fl.layout = ncol, nlev, layout
        if isinstance(fields, (list, tuple)):
            assert len(fields) == L.NSTATE
            ld = None
            for f, t in enumerate(fields):
                if t is None:
                    continue
                assert t.dim() == 2 and t.dtype == torch.float64 and t.is_cuda
                assert t.stride(1) == 1 and ld in (None, t.stride(0))
                ld = t.stride(0)
                fl.field[f] = t.data_ptr()
            fl.ldc = ld or 0
        else:  # a (possibly padded) [52][slow][fast] tensor
            assert fields.dim() == 3 and fields.shape[0] == L.NSTATE, fields.shape
            assert fields.dtype == torch.float64 and fields.is_cuda and fields.stride(2) == 1
            for f in range(L.NSTATE):
                fl.field[f] = fields.data_ptr() + f * fields.stride(0) * 8
            fl.ldc = fields.stride(1)
        return fl

    def tracers_to_fields(self, state_q, qqcw, fields, layout=_abi.LAYOUT_LEV_COL, stream=None):
        """state_q / qqcw (ncol, nlev, 40) -> the 52 state planes
        (mam4x_tracers_to_fields; inject_stateq/qqcw_to_prognostics)"""
        ncol, nlev = state_q.shape[:2]
        tr, fl = self._tracers(state_q, qqcw), self._fields(fields, ncol, nlev, layout)
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_tracers_to_fields(self.h, ctypes.byref(tr), ctypes.byref(fl),
                                              ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_tracers_to_fields")

    def fields_to_tracers(self, fields, state_q, qqcw, mode=_abi.PACK_STATE,
                          layout=_abi.LAYOUT_LEV_COL, stream=None):
        """the 52 state (or tendency) planes -> state_q / qqcw
        (mam4x_fields_to_tracers; extract_stateq/qqcw_from_prognostics, or
        extract_ptend_from_tendencies with mode=PACK_TENDENCY)"""
        ncol, nlev = state_q.shape[:2]
        tr, fl = self._tracers(state_q, qqcw), self._fields(fields, ncol, nlev, layout)
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_fields_to_tracers(self.h, ctypes.byref(fl), ctypes.byref(tr), mode,
                                              ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_fields_to_tracers")

    def compute_gas_chemistry(self, dt, temperature, photolysis_rate, extfrc, invariants,
                              het_rates, qq, layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """Gas-phase chemistry (mam4x_compute_gas_chemistry).  temperature /
        photolysis_rate: planes; extfrc [9][...], invariants [8][...],
        het_rates / qq [31][...] tensors (qq advanced in place); None
        extfrc / het_rates = 0.  live: see _extents."""
        a = _abi.GasChemArgs()
        t = temperature
        a.ncol, a.nlev, a.ldc = self._extents(t, layout, live)
        a.layout = layout
        a.temperature = t.data_ptr()
        assert photolysis_rate.shape == t.shape and photolysis_rate.is_contiguous()
        a.photolysis_rate = photolysis_rate.data_ptr()
        for name, arr, n in (("extfrc", extfrc, _abi.EXTCNT), ("invariants", invariants, _abi.NFS),
                             ("het_rates", het_rates, _abi.GAS_PCNST), ("qq", qq, _abi.GAS_PCNST)):
            if arr is None:
                continue
            for i, p in enumerate(_plane_ptrs(arr, n)):
                getattr(a, name)[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_gas_chemistry(self.h, dt, ctypes.byref(a),
                                                  ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_gas_chemistry")

    @staticmethod
    def _extents(t, layout, live=None):
        """(ncol, nlev, ldc) of a contiguous float64 CUDA plane.  The plane's
        fastest dimension is the leading dimension ldc; live (default: all of
        it) is how many of its entries the domain uses -- ncol for
        LAYOUT_LEV_COL, nlev for LAYOUT_COL_LEV -- as with planes a host
        dimensions for its largest chunk (pcols) and fills up to ncol."""
        assert t.dim() == 2 and t.dtype == torch.float64 and t.is_cuda and t.is_contiguous()
        slow, ldc = t.shape
        fast = ldc if live is None else int(live)
        assert 0 <= fast <= ldc, (fast, ldc)
        if layout == _abi.LAYOUT_LEV_COL:
            return fast, slow, ldc
        return slow, fast, ldc

    def compute_invariants(self, temperature, pressure, qv, cnst_offline, invariants,
                           layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """Invariant densities (mam4x_compute_invariants, setinv): planes T,
        p, qv; cnst_offline [4][...] or None; invariants [8][...] output.
        live: see _extents."""
        a = _abi.SetinvArgs()
        a.ncol, a.nlev, a.ldc = self._extents(temperature, layout, live)
        a.layout = layout
        for name, t in (("temperature", temperature), ("pressure", pressure),
                        ("vapor_mixing_ratio", qv)):
            assert t.shape == temperature.shape and t.is_contiguous()
            setattr(a, name, t.data_ptr())
        if cnst_offline is not None:
            for i, p in enumerate(_plane_ptrs(cnst_offline, 4)):
                a.cnst_offline[i] = p
        for i, p in enumerate(_plane_ptrs(invariants, _abi.NFS)):
            a.invariants[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_invariants(self.h, ctypes.byref(a),
                                               ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_invariants")

    def compute_linoz(self, dt, temperature, pressure, pdel, climatology, zenith_angle,
                      latitude, ltrop, o3, conf, column_above_top=None, o3_column=None,
                      do3_linoz=None, do3_linoz_psc=None, ss_o3=None, o3_sfcsink=None,
                      layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """Stratospheric ozone (mam4x_compute_linoz): planes T, p, pdel, o3
        (advanced in place) and the optional output planes; climatology
        [8][...]; per column zenith_angle, latitude, ltrop (int32),
        column_above_top and o3_sfcsink; conf: dict of chlorine_loading,
        psc_T, o3_sfc, o3_tau, o3_lbl.  live: see _extents."""
        a = _abi.LinozArgs()
        a.ncol, a.nlev, a.ldc = self._extents(temperature, layout, live)
        a.layout = layout
        for name, t in (("temperature", temperature), ("pressure", pressure), ("pdel", pdel),
                        ("o3", o3), ("o3_column", o3_column), ("do3_linoz", do3_linoz),
                        ("do3_linoz_psc", do3_linoz_psc), ("ss_o3", ss_o3)):
            if t is None:
                continue
            assert t.shape == temperature.shape and t.dtype == torch.float64 and t.is_contiguous()
            setattr(a, name, t.data_ptr())
        for i, p in enumerate(_plane_ptrs(climatology, _abi.NLINOZ)):
            a.climatology[i] = p
        for name, t, dt_ in (("zenith_angle", zenith_angle, torch.float64),
                             ("latitude", latitude, torch.float64), ("ltrop", ltrop, torch.int32),
                             ("column_above_top", column_above_top, torch.float64),
                             ("o3_sfcsink", o3_sfcsink, torch.float64)):
            if t is None:
                continue
            assert t.shape == (a.ncol,) and t.dtype == dt_ and t.is_cuda and t.is_contiguous()
            setattr(a, name, t.data_ptr())
        for k in ("chlorine_loading", "psc_T", "o3_sfc", "o3_tau"):
            setattr(a, k, float(conf[k]))
        a.o3_lbl = int(conf["o3_lbl"])
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_linoz(self.h, dt, ctypes.byref(a),
                                          ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_linoz")

    def compute_photolysis(self, table, planes, zenith_angle, surface_albedo, photo,
                           esfact=1.0, work=None, layout=_abi.LAYOUT_LEV_COL, stream=None,
                           live=None):
        """Photolysis rates (mam4x_compute_photolysis, table_photo).  table:
        dict of the mam4x_photo_table arrays as CUDA float64 tensors
        (rsf_tab [nw, nump, numsza, numcolo3, numalb], xsqy [numj, nw, nt,
        np_xs], the axes) and optionally pht_alias_mult, lng_index; planes:
        dict of pressure, pdel, temperature, o3_column, liquid_water,
        cloud_fraction tensors; zenith_angle, surface_albedo [ncol]; photo:
        plane updated in place; work: scratch of mam4x_photo_work_len doubles
        (allocated when None; returned for reuse).  live: see _extents."""
        tb = _abi.PhotoTable()
        for k in _abi.PHOTO_TABLE_ARRAYS:
            v = table[k]
            assert v.dtype == torch.float64 and v.is_cuda and v.is_contiguous(), k
            setattr(tb, k, v.data_ptr())
        tb.numj, tb.nw, tb.nt, tb.np_xs = table["xsqy"].shape
        _, tb.nump, tb.numsza, tb.numcolo3, tb.numalb = table["rsf_tab"].shape
        tb.pht_alias_mult = float(table.get("pht_alias_mult", 1.0))
        tb.lng_index = int(table.get("lng_index", 0))
        a = _abi.PhotoArgs()
        a.ncol, a.nlev, a.ldc = self._extents(photo, layout, live)
        a.layout = layout
        for k in ("pressure", "pdel", "temperature", "o3_column", "liquid_water",
                  "cloud_fraction"):
            v = planes[k]
            assert v.shape == photo.shape and v.dtype == torch.float64 and v.is_contiguous(), k
            setattr(a, k, v.data_ptr())
        for name, v in (("zenith_angle", zenith_angle), ("surface_albedo", surface_albedo)):
            assert v.shape == (a.ncol,) and v.dtype == torch.float64 and v.is_contiguous()
            setattr(a, name, v.data_ptr())
        a.esfact = float(esfact)
        a.photo = photo.data_ptr()
        need = self.lib.mam4x_photo_work_len(ctypes.byref(tb), a.ncol, a.nlev, a.ldc, layout)
        assert need >= 0
        if work is None:
            work = torch.empty(need, dtype=torch.float64, device=photo.device)
        assert work.dtype == torch.float64 and work.is_contiguous() and work.numel() >= need
        a.work = work.data_ptr()
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_photolysis(self.h, ctypes.byref(tb), ctypes.byref(a),
                                               ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_photolysis")
        return work

    def compute_washout(self, dt, planes, latitude, phis, het_rates,
                        layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """Washout rates (mam4x_compute_washout, sethet): planes: dict of
        pressure, zmid, temperature, cmfdqr, nrain, nevapr, air_density,
        h2o2 tensors; latitude, phis [ncol]; het_rates [31][...] output.
        live: see _extents."""
        a = _abi.SethetArgs()
        t = planes["temperature"]
        a.ncol, a.nlev, a.ldc = self._extents(t, layout, live)
        a.layout = layout
        for k in ("pressure", "zmid", "temperature", "cmfdqr", "nrain", "nevapr",
                  "air_density", "h2o2"):
            v = planes[k]
            assert v.shape == t.shape and v.dtype == torch.float64 and v.is_contiguous(), k
            setattr(a, k, v.data_ptr())
        for name, v in (("latitude", latitude), ("phis", phis)):
            assert v.shape == (a.ncol,) and v.dtype == torch.float64 and v.is_contiguous()
            setattr(a, name, v.data_ptr())
        for i, p in enumerate(_plane_ptrs(het_rates, _abi.GAS_PCNST)):
            a.het_rates[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_washout(self.h, dt, ctypes.byref(a),
                                            ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_washout")

    def compute_forcing(self, sectors, extfrc, frc_ndx=None, file_alt_data=None,
                        layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """External forcing (mam4x_compute_forcing, setext): sectors: 9 lists
        (one per forced species) of up to 4 planes; extfrc [9][...] output;
        frc_ndx: 1-based destination entries (default 1..9); file_alt_data:
        per species, data stored bottom-up.  live: see _extents."""
        a = _abi.SetextArgs()
        t = extfrc[0]
        a.ncol, a.nlev, a.ldc = self._extents(t, layout, live)
        a.layout = layout
        assert len(sectors) == _abi.EXTCNT
        for m in range(_abi.EXTCNT):
            a.frc_ndx[m] = frc_ndx[m] if frc_ndx is not None else m + 1
            a.file_alt_data[m] = int(bool(file_alt_data[m])) if file_alt_data is not None else 0
            a.nsectors[m] = len(sectors[m])
            for i, v in enumerate(sectors[m]):
                assert v.shape == t.shape and v.dtype == torch.float64 and v.is_contiguous()
                a.sectors[m][i] = v.data_ptr()
        for i, p in enumerate(_plane_ptrs(extfrc, _abi.EXTCNT)):
            a.extfrc[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_forcing(self.h, ctypes.byref(a),
                                            ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_forcing")

    def compute_setsox(self, dt, met, qcw, qin, dqdt_aqso4=None, dqdt_aqh2so4=None,
                       layout=_abi.LAYOUT_LEV_COL, stream=None, live=None):
        """Aqueous sulfur chemistry (mam4x_compute_setsox).  met: dict of
        planes pressure, temperature, liquid_water, cloud_fraction,
        droplet_number, air_number_density; qcw / qin: [31][nlev][ncol]
        tensors (gas_pcnst order) updated in place; dqdt_*: optional
        [31][nlev][ncol] outputs ([ncol][nlev] planes for LAYOUT_COL_LEV).
        live: see _extents."""
        a = _abi.SetsoxArgs()
        t = met["temperature"]
        a.ncol, a.nlev, a.ldc = self._extents(t, layout, live)
        a.layout = layout
        for k in ("pressure", "temperature", "liquid_water", "cloud_fraction", "droplet_number",
                  "air_number_density"):
            v = met[k]
            assert v.shape == t.shape and v.dtype == torch.float64 and v.is_contiguous()
            setattr(a, k, v.data_ptr())
        for name, arr in (("qcw", qcw), ("qin", qin), ("dqdt_aqso4", dqdt_aqso4),
                          ("dqdt_aqh2so4", dqdt_aqh2so4)):
            if arr is None:
                continue
            for i, p in enumerate(_plane_ptrs(arr, _abi.GAS_PCNST)):
                getattr(a, name)[i] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_setsox(self.h, dt, ctypes.byref(a),
                                           ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_setsox")

    def make_outputs(self, tend, dgn_i=None, dgn_c=None, state_out=None):
        out = _abi.Outputs()
        for i, p in enumerate(_plane_ptrs(tend, L.NSTATE)):
            out.tendencies[i] = p
        if dgn_i is not None:
            for m, p in enumerate(_plane_ptrs(dgn_i, 4)):
                out.dgncur_i[m] = p
        if dgn_c is not None:
            for m, p in enumerate(_plane_ptrs(dgn_c, 4)):
                out.dgncur_c[m] = p
        if state_out is not None:
            for i, p in enumerate(_plane_ptrs(state_out, L.NSTATE)):
                out.state_out[i] = p
        return out

    def batch(self, scenes, dt, tends, dgn_i=None, dgn_c=None, state_out=None,
              layout=_abi.LAYOUT_LEV_COL, budgets=None):
        """Plan one step of many scenes (lists of scene dicts and output
        tensors, None entries allowed in the optional lists; budgets: float64
        [NBUDGET][ncol] per scene); see Batch."""
        return Batch(self, scenes, dt, tends, dgn_i, dgn_c, state_out, layout, budgets)

    def compute_tangent(self, scene, dt, tend, dtend, dstate=None, dtemperature=None,
                        dpressure=None, dqv=None, dgn_i=None, dgn_c=None, t=0.0, stream=None,
                        layout=_abi.LAYOUT_LEV_COL, precldchem=None):
        """Tendencies and their derivative along (dstate, dT, dp, dqv) (None =
        zero direction), through mam4x_compute_tangent.  precldchem: the
        optional pre-cloud-chemistry state (held constant)."""
        inp = self.make_inputs(scene, precldchem, None, layout)
        out = _abi.Outputs()
        for i, p in enumerate(_plane_ptrs(tend, L.NSTATE)):
            out.tendencies[i] = p
        if dgn_i is not None:
            for m, p in enumerate(_plane_ptrs(dgn_i, 4)):
                out.dgncur_i[m] = p
        if dgn_c is not None:
            for m, p in enumerate(_plane_ptrs(dgn_c, 4)):
                out.dgncur_c[m] = p
        tg = _abi.Tangent()
        for i, p in enumerate(_plane_ptrs(dtend, L.NSTATE)):
            tg.tendencies[i] = p
        if dstate is not None:
            for i, p in enumerate(_plane_ptrs(dstate, L.NSTATE)):
                tg.state[i] = p
        for name, tt in (("temperature", dtemperature), ("pressure", dpressure),
                         ("vapor_mixing_ratio", dqv)):
            if tt is not None:
                assert tt.dtype == torch.float64 and tt.is_cuda and tt.is_contiguous()
                setattr(tg, name, tt.data_ptr())
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_tangent(self.h, t, dt, ctypes.byref(inp), ctypes.byref(tg),
                                            ctypes.byref(out), ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_tangent")

    def compute_adjoint(self, scene, dt, tbar, state_bar=None, temperature_bar=None,
                        pressure_bar=None, qv_bar=None, t=0.0, stream=None,
                        layout=_abi.LAYOUT_LEV_COL, precldchem=None):
        """Per-cell vector-Jacobian product through mam4x_compute_adjoint:
        tbar ([52, ...] cotangent of the tendencies, or a list of 52 planes
        with None for zero planes) in; the gradients of
        sum(tbar * tendencies) with respect to the state, T, p and qv written
        to the given tensors (None = not formed).  precldchem: the optional
        pre-cloud-chemistry state (held constant)."""
        inp = self.make_inputs(scene, precldchem, None, layout)
        adj = _abi.Adjoint()
        if isinstance(tbar, (list, tuple)):  # one plane or None per field
            assert len(tbar) == L.NSTATE
            for i, pl in enumerate(tbar):
                if pl is not None:
                    assert pl.dtype == torch.float64 and pl.is_cuda and pl.is_contiguous()
                    adj.tendencies[i] = pl.data_ptr()
        else:
            for i, p in enumerate(_plane_ptrs(tbar, L.NSTATE)):
                adj.tendencies[i] = p
        if state_bar is not None:
            for i, p in enumerate(_plane_ptrs(state_bar, L.NSTATE)):
                adj.state[i] = p
        for name, tt in (("temperature", temperature_bar), ("pressure", pressure_bar),
                         ("vapor_mixing_ratio", qv_bar)):
            if tt is not None:
                assert tt.dtype == torch.float64 and tt.is_cuda and tt.is_contiguous()
                setattr(adj, name, tt.data_ptr())
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_adjoint(self.h, t, dt, ctypes.byref(inp), ctypes.byref(adj),
                                            ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_adjoint")

    def compute_water_uptake(self, scene, dgncur_awet, qaerwat, wetdens=None, dgncur_a=None,
                             stream=None, layout=_abi.LAYOUT_LEV_COL):
        """Aerosol water uptake through mam4x_compute_water_uptake: writes the
        wet diameters, aerosol water and (optionally) wet density, [4, ...]
        each, from the scene's interstitial masses, T, p, qv, cloud fraction
        and dry diameters (dgncur_a overrides the scene's)."""
        if dgncur_a is not None:
            scene = dict(scene)
            scene["dgncur_a"] = dgncur_a
        inp = self.make_inputs(scene, None, None, layout)
        out = _abi.WaterOutputs()
        for m, p in enumerate(_plane_ptrs(dgncur_awet, 4)):
            out.dgncur_awet[m] = p
        for m, p in enumerate(_plane_ptrs(qaerwat, 4)):
            out.qaerwat[m] = p
        if wetdens is not None:
            for m, p in enumerate(_plane_ptrs(wetdens, 4)):
                out.wetdens[m] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_water_uptake(self.h, ctypes.byref(inp), ctypes.byref(out),
                                                 ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_water_uptake")

    def compute_wet_size(self, scene, dt, dgncur_i, dgncur_awet, qaerwat, dgncur_c=None,
                         wetdens=None, tend=None, stream=None, layout=_abi.LAYOUT_LEV_COL):
        """The wet-size producer through mam4x_compute_wet_size: E3SM-layout
        calcsize (dry diameters, optional tendencies [52, ...]) then water
        uptake on those diameters; outputs [4, ...] each."""
        inp = self.make_inputs(scene, None, None, layout)
        out = _abi.WetSizeOutputs()
        for key, t in (("dgncur_i", dgncur_i), ("dgncur_c", dgncur_c),
                       ("dgncur_awet", dgncur_awet), ("qaerwat", qaerwat), ("wetdens", wetdens)):
            if t is not None:
                arr = getattr(out, key)
                for m, p in enumerate(_plane_ptrs(t, 4)):
                    arr[m] = p
        if tend is not None:
            for f, p in enumerate(_plane_ptrs(tend, L.NSTATE)):
                out.tendencies[f] = p
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_compute_wet_size(self.h, dt, ctypes.byref(inp), ctypes.byref(out),
                                             ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_compute_wet_size")

    def validate(self, scene, stream=None):
        inp = self.make_inputs(scene)
        v = ctypes.c_longlong(0)
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.mam4x_validate(self.h, ctypes.byref(inp), ctypes.byref(v),
                                     ctypes.c_void_p(s.cuda_stream))
        self._check(rc, "mam4x_validate")
        return int(v.value)

    def enable_stats(self, on=True):
        self._check(self.lib.mam4x_enable_stats(self.h, int(on)), "mam4x_enable_stats")

    def last_stats(self, stream=None):
        v = ctypes.c_longlong(0)
        s = stream if stream is not None else torch.cuda.current_stream()
        self._check(self.lib.mam4x_last_stats(self.h, ctypes.byref(v),
                                              ctypes.c_void_p(s.cuda_stream)), "mam4x_last_stats")
        return int(v.value)


    def last_phase_times(self, stream=None):
        """{phase: milliseconds} summed over the calls since the last read,
        and the call count (stats must be enabled)."""
        ms = (ctypes.c_double * _abi.NPHASE)()
        n = ctypes.c_longlong(0)
        s = stream if stream is not None else torch.cuda.current_stream()
        self._check(self.lib.mam4x_last_phase_times(self.h, ms, ctypes.byref(n),
                                                    ctypes.c_void_p(s.cuda_stream)),
                    "mam4x_last_phase_times")
        return dict(zip(_abi.PHASE_NAMES, list(ms))), int(n.value)


class Batch:
    """A planned multi-scene step (mam4x_batch_*).  The tensors' storage is
    bound at creation; run() launches kernels only and can be captured in a
    CUDA graph (torch.cuda.graph) and replayed."""

    def __init__(self, chain, scenes, dt, tends, dgn_i, dgn_c, state_out, layout, budgets=None):
        n = len(scenes)
        pick = lambda lst, i: None if lst is None else lst[i]
        ins = (_abi.Inputs * n)()
        outs = (_abi.Outputs * n)()
        for i, sc in enumerate(scenes):
            ins[i] = chain.make_inputs(sc, None, None, layout)
            outs[i] = chain.make_outputs(tends[i], pick(dgn_i, i), pick(dgn_c, i),
                                         pick(state_out, i))
            bud = pick(budgets, i)
            if bud is not None:
                assert bud.dtype == torch.float64 and bud.is_cuda and bud.is_contiguous()
                assert bud.shape == (_abi.NBUDGET, ins[i].ncol), bud.shape
                outs[i].column_budgets = bud.data_ptr()
        self._keep = (scenes, tends, dgn_i, dgn_c, state_out, budgets)
        self.chain, self.dt, self.nscene = chain, dt, n
        b = ctypes.c_void_p()
        chain._check(chain.lib.mam4x_batch_create(chain.h, n, ins, outs, dt, ctypes.byref(b)),
                     "mam4x_batch_create")
        self.b = b

    def run(self, t=0.0, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream()
        self.chain._check(self.chain.lib.mam4x_batch_run(self.b, t, self.dt,
                                                          ctypes.c_void_p(s.cuda_stream)),
                          "mam4x_batch_run")

    def close(self):
        if getattr(self, "b", None):
            self.chain.lib.mam4x_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_to_device(scene, device="cuda"):
    """numpy scene dict -> dict of contiguous float64 CUDA tensors."""
    return {k: torch.as_tensor(v, dtype=torch.float64).contiguous().to(device)
            for k, v in scene.items()}
