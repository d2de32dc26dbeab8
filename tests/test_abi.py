"""C ABI checks that need no GPU: the library loads, exports every function
include/mam4x.h declares, and the ctypes mirror has the C struct layouts."""
import ctypes
import os
import re
import subprocess

import pytest

from mam4xx_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "mam4x.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:long\s+)?\w+\s*\*?\s*(mam4x_\w+)\s*\(",
                                 src, re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "mam4x_compute_tendencies" in names and "mam4x_create" in names
    assert set(names) == set(_abi.SIGNATURES), set(names) ^ set(_abi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mam4x_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_c(tmp_path):
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "mam4x.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mam4x_config),'
                    ' sizeof(mam4x_inputs), sizeof(mam4x_outputs),'
                    ' offsetof(mam4x_inputs, precldchem), offsetof(mam4x_outputs, state_out),'
                    ' offsetof(mam4x_outputs, process_tendencies),'
                    ' offsetof(mam4x_outputs, process_stride));'
                    'printf("%zu %zu %zu\\n", sizeof(mam4x_tracers), sizeof(mam4x_fields),'
                    ' offsetof(mam4x_fields, field));'
                    'printf("%zu %zu\\n", sizeof(mam4x_setsox_args),'
                    ' offsetof(mam4x_setsox_args, dqdt_aqh2so4));'
                    'printf("%zu %zu\\n", sizeof(mam4x_gas_chem_args),'
                    ' offsetof(mam4x_gas_chem_args, qq));'
                    'printf("%zu %zu\\n", sizeof(mam4x_setinv_args),'
                    ' offsetof(mam4x_setinv_args, invariants));'
                    'printf("%zu %zu %zu %zu\\n", sizeof(mam4x_linoz_args),'
                    ' offsetof(mam4x_linoz_args, chlorine_loading),'
                    ' offsetof(mam4x_linoz_args, o3_lbl), offsetof(mam4x_linoz_args, o3_sfcsink));'
                    'printf("%zu %zu\\n", sizeof(mam4x_sethet_args),'
                    ' offsetof(mam4x_sethet_args, het_rates));'
                    'printf("%zu %zu %zu\\n", sizeof(mam4x_setext_args),'
                    ' offsetof(mam4x_setext_args, sectors), offsetof(mam4x_setext_args, extfrc));'
                    'printf("%zu %zu %zu\\n", sizeof(mam4x_photo_table),'
                    ' offsetof(mam4x_photo_table, rsf_tab), offsetof(mam4x_photo_table, lng_index));'
                    'printf("%zu %zu %zu\\n", sizeof(mam4x_photo_args),'
                    ' offsetof(mam4x_photo_args, esfact), offsetof(mam4x_photo_args, work));'
                    'printf("%zu %zu %d %d %d\\n", offsetof(mam4x_inputs, hydrostatic_dp),'
                    ' offsetof(mam4x_outputs, column_budgets), MAM4X_NBUDGET,'
                    ' MAM4X_BUDGET_SULFUR, MAM4X_BUDGET_ORGANIC);'
                    'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(prog)],
                   check=True)
    c = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()]
    py = [ctypes.sizeof(_abi.Config), ctypes.sizeof(_abi.Inputs), ctypes.sizeof(_abi.Outputs),
          _abi.Inputs.precldchem.offset, _abi.Outputs.state_out.offset,
          _abi.Outputs.process_tendencies.offset, _abi.Outputs.process_stride.offset,
          ctypes.sizeof(_abi.Tracers), ctypes.sizeof(_abi.Fields), _abi.Fields.field.offset,
          ctypes.sizeof(_abi.SetsoxArgs), _abi.SetsoxArgs.dqdt_aqh2so4.offset,
          ctypes.sizeof(_abi.GasChemArgs), _abi.GasChemArgs.qq.offset,
          ctypes.sizeof(_abi.SetinvArgs), _abi.SetinvArgs.invariants.offset,
          ctypes.sizeof(_abi.LinozArgs), _abi.LinozArgs.chlorine_loading.offset,
          _abi.LinozArgs.o3_lbl.offset, _abi.LinozArgs.o3_sfcsink.offset,
          ctypes.sizeof(_abi.SethetArgs), _abi.SethetArgs.het_rates.offset,
          ctypes.sizeof(_abi.SetextArgs), _abi.SetextArgs.sectors.offset,
          _abi.SetextArgs.extfrc.offset,
          ctypes.sizeof(_abi.PhotoTable), _abi.PhotoTable.rsf_tab.offset,
          _abi.PhotoTable.lng_index.offset,
          ctypes.sizeof(_abi.PhotoArgs), _abi.PhotoArgs.esfact.offset, _abi.PhotoArgs.work.offset,
          _abi.Inputs.hydrostatic_dp.offset, _abi.Outputs.column_budgets.offset, _abi.NBUDGET,
          _abi.BUDGET_SULFUR, _abi.BUDGET_ORGANIC]
    assert c == py


def test_abi_version_and_default_config():
    lib = _abi.load()
    assert lib.mam4x_abi_version() == _abi.ABI_VERSION == 11
    cfg = _abi.Config()
    assert lib.mam4x_default_config(ctypes.byref(cfg)) == _abi.OK
    assert (cfg.do_cond, cfg.do_rename, cfg.do_newnuc, cfg.do_coag) == (1, 1, 1, 1)
    assert cfg.n_so4_monolayers_pcage == 8 and cfg.precision == _abi.PREC_FP64
    assert cfg.ntsubstep == 1 and cfg.levels_per_wave == 1
    assert (cfg.soa_chunk, cfg.soa_short_chunk, cfg.soa_short_bucket) == (0, 0, 0)
    assert lib.mam4x_default_config(None) == _abi.ERR_ARG


def test_null_handle_is_an_error_not_a_crash():
    lib = _abi.load()
    assert lib.mam4x_destroy(None) == _abi.ERR_ARG
    assert lib.mam4x_set_precision(None, 0) == _abi.ERR_ARG
    assert lib.mam4x_init(None, None) == _abi.ERR_ARG
    assert lib.mam4x_compute_tendencies(None, 0.0, 1.0, None, None, None) == _abi.ERR_ARG
    assert lib.mam4x_last_error(None) == b"NULL handle"


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_abi.LibraryMissing):
        _abi.load(str(tmp_path / "nope.so"))
