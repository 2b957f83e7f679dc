"""The C-ABI library loads and exports every entry point include/hora_hip.h declares (no GPU needed)."""
import ctypes as C
import os
import re

from ptctcl_amd import hip

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "hora_hip.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(hora_hip_[a-z_]+)\s*\(", txt)))


def test_header_declares_expected():
    assert set(declared()) == set(hip.EXPORTS)


def test_library_exports_all():
    L = hip.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_abi_version_without_device():
    assert hip.lib().hora_hip_abi_version() == 1
    assert b"gfx950" in hip.lib().hora_hip_build_info()


def test_struct_layout_matches_c(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's sizes and field offsets."""
    import shutil
    import subprocess
    from ptctcl_amd.network import hora_net_enc, hora_reach_stats, hora_target
    from ptctcl_amd.hip import hora_reach_profile
    from ptctcl_amd.liveness import hora_dpred, hora_gfp_graph, hora_live_graph, hora_live_query, hora_live_stats
    from ptctcl_amd.serve import hora_serve_job_result
    structs = (hora_net_enc, hora_target, hora_reach_stats, hora_reach_profile, hora_dpred, hora_live_query,
               hora_live_stats, hora_live_graph, hora_gfp_graph, hora_serve_job_result)
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void){"]
    for S in structs:
        lines.append(f'printf("{S.__name__} %zu\\n", sizeof({S.__name__}));')
        for f, _ in S._fields_:
            lines.append(f'printf("{S.__name__}.{f} %zu\\n", offsetof({S.__name__}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for S in structs:
        assert int(got[S.__name__]) == C.sizeof(S)
        for f, _ in S._fields_:
            assert int(got[f"{S.__name__}.{f}"]) == getattr(S, f).offset, (S.__name__, f)
