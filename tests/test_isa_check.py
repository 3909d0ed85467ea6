"""CPU: the build-time ISA check flags an address computation that reads an
SGPR left undefined on one path, and passes the same code once every path
defines it."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "ch_swe_amd"))
import isa_check  # noqa: E402

# two paths into .LBB0_3: the first writes s[28:29], the second leaves it as
# an implicit-def; .LBB0_3 stores through it
BAD = """
_Z4kernv: ; @_Z4kernv
; %bb.0:
\ts_load_dwordx4 s[8:11], s[0:1], 0x0
\ts_cmp_eq_u32 s2, 3
\ts_cbranch_scc1 .LBB0_2
; %bb.1:
\ts_mov_b64 s[28:29], s[8:9]
\ts_branch .LBB0_3
.LBB0_2:
                                        ; implicit-def: $sgpr28_sgpr29
\ts_nop 0
.LBB0_3:
\tv_lshl_add_u64 v[4:5], v[2:3], 3, s[28:29]
\tglobal_store_dwordx2 v[4:5], v[6:7], off
\ts_endpgm
.Lfunc_end0:
"""
GOOD = BAD.replace("""                                        ; implicit-def: $sgpr28_sgpr29
\ts_nop 0""", "\ts_mov_b64 s[28:29], s[10:11]")


def _check(text, tmp_path):
    p = tmp_path / "k.s"
    p.write_text(text)
    return isa_check.check_file(str(p))


def test_flags_store_base_undefined_on_one_path(tmp_path):
    found = _check(BAD, tmp_path)
    assert list(found) == ["_Z4kernv"]
    assert {r[3] for r in found["_Z4kernv"]} == {28, 29}


def test_passes_when_every_path_defines_the_base(tmp_path):
    assert _check(GOOD, tmp_path) == {}


def test_built_library_assembly_is_clean():
    """The assembly written by the last build (if any) has no findings."""
    asm = os.path.join(HERE, "..", "ch_swe_amd", "_build", "swe_lib_gfx950.s")
    if not os.path.exists(asm):
        import pytest

        pytest.skip("library not built in this tree")
    assert isa_check.check_file(asm) == {}


# .LBB0_3 (the store) is reached from .LBB0_2, where s[28:29] is defined, and
# from .LBB0_1 only through a saveexec of a mask set to zero just before: that
# path always takes the execz branch to .LBB0_4 and never stores
ZERO_EXEC = """
_Z4kernv: ; @_Z4kernv
; %bb.0:
\ts_load_dwordx4 s[8:11], s[0:1], 0x0
\ts_cmp_eq_u32 s2, 3
\ts_cbranch_scc1 .LBB0_2
; %bb.1:
\ts_mov_b64 s[14:15], 0
                                        ; implicit-def: $sgpr28_sgpr29
\ts_branch .LBB0_5
.LBB0_2:
\ts_mov_b64 s[28:29], s[8:9]
\ts_mov_b64 s[14:15], exec
.LBB0_5:
\ts_and_saveexec_b64 s[0:1], s[14:15]
\ts_cbranch_execz .LBB0_4
.LBB0_3:
\tv_lshl_add_u64 v[4:5], v[2:3], 3, s[28:29]
\tglobal_store_dwordx2 v[4:5], v[6:7], off
.LBB0_4:
\ts_endpgm
.Lfunc_end0:
"""


def test_branch_on_zeroed_exec_mask_is_not_a_path(tmp_path):
    assert _check(ZERO_EXEC, tmp_path) == {}


def test_nonzero_mask_keeps_the_path(tmp_path):
    live = ZERO_EXEC.replace("\ts_mov_b64 s[14:15], 0\n", "\ts_mov_b64 s[14:15], exec\n")
    found = _check(live, tmp_path)
    assert {r[3] for r in found["_Z4kernv"]} == {28, 29}


def test_zeroed_mask_with_scalar_work_before_the_branch(tmp_path):
    """Instructions between the saveexec and the execz branch that do not
    write exec keep the branch taken."""
    txt = ZERO_EXEC.replace("\ts_and_saveexec_b64 s[0:1], s[14:15]\n",
                            "\ts_and_saveexec_b64 s[0:1], s[14:15]\n\tv_readlane_b32 s4, v7, 4\n")
    assert _check(txt, tmp_path) == {}
    bad = txt.replace("\tv_readlane_b32 s4, v7, 4\n", "\ts_mov_b64 exec, s[0:1]\n")
    assert {r[3] for r in _check(bad, tmp_path)["_Z4kernv"]} == {28, 29}


# the store block is entered by execnz on a lane mask that only the block
# defining the store base sets; on the path that skips that block the mask is
# still the zero it was given, so the execnz branch is not taken
EXECNZ = """
_Z4kernv: ; @_Z4kernv
; %bb.0:
\ts_load_dwordx4 s[8:11], s[0:1], 0x0
\ts_mov_b64 s[20:21], 0
\tv_cmp_gt_u32_e32 vcc, 7, v0
                                        ; implicit-def: $sgpr18_sgpr19
\ts_and_saveexec_b64 s[2:3], vcc
\ts_cbranch_execz .LBB0_2
; %bb.1:
\ts_mov_b64 s[18:19], s[8:9]
\ts_or_b64 s[20:21], s[20:21], exec
.LBB0_2:
\ts_or_b64 exec, exec, s[2:3]
\ts_and_saveexec_b64 s[0:1], s[20:21]
\ts_cbranch_execnz .LBB0_3
; %bb.4:
\ts_endpgm
.LBB0_3:
\tv_lshl_add_u64 v[4:5], s[18:19], 3, v[2:3]
\tglobal_store_dwordx2 v[4:5], v[6:7], off
\ts_endpgm
.Lfunc_end0:
"""


def test_execnz_on_a_mask_still_zero_is_not_a_path(tmp_path):
    assert _check(EXECNZ, tmp_path) == {}
    # a mask that is not known to be zero on the skip path keeps the finding
    live = EXECNZ.replace("\ts_mov_b64 s[20:21], 0\n", "\ts_mov_b64 s[20:21], exec\n")
    assert {r[3] for r in _check(live, tmp_path)["_Z4kernv"]} == {18, 19}


def test_out_of_line_device_function_is_refused(tmp_path):
    """A device function the compiler did not inline is reported: the
    analysis assumes a kernel's entry state."""
    text = GOOD + """
_Z6helperPf: ; @_Z6helperPf
; %bb.0:
\ts_waitcnt vmcnt(0)
\ts_setpc_b64 s[30:31]
.Lfunc_end1:
\t.amdhsa_kernel _Z4kernv
\t.end_amdhsa_kernel
"""
    found = _check(text, tmp_path)
    assert list(found) == ["_Z6helperPf"]
    assert found["_Z6helperPf"][0][1] == "out-of-line device function"
