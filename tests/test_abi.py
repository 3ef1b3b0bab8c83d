"""The C-ABI libraries load and export every function include/*.h declares
(no GPU needed: nothing is launched)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(velo_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("velo_hlt.h", "libvelo_hlt.so"), ("velo_gen.h", "libvelo_gen.so"),
                                        ("velo_sequence.h", "libvelo_hlt.so"), ("velo_checker.h", "libvelo_hlt.so"),
                                        ("velo_ghost.h", "libvelo_hlt.so"), ("velo_mdf.h", "libvelo_hlt.so"),
                                        ("velo_pv.h", "libvelo_hlt.so")])
def test_library_exports_declared_symbols(header, lib):
    import ctypes
    names = declared(header)
    assert names, header
    L = ctypes.CDLL(str(ROOT / "cuda_hlt_amd" / "lib" / lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_structs_match_header_layout():
    """Field order of the ctypes mirrors equals the header's structs."""
    from cuda_hlt_amd import native as nat
    text = (ROOT / "include" / "velo_hlt.h").read_text() + (ROOT / "include" / "velo_pv.h").read_text()
    for m in re.finditer(r"typedef struct \{([^{}]*)\} (velo_\w+_args);", text):
        body, name = m.groups()
        fields = re.findall(r"\b(\w+);", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
        launcher = name[:-5] + "_launch"
        st = nat.LAUNCHERS[launcher]
        assert [f for f, _ in st._fields_] == fields, name


def test_error_strings():
    from cuda_hlt_amd import native as nat
    assert nat.lib().velo_error_string(0)
    assert b"gfx950" in nat.lib().velo_hlt_version()


def _header_structs(header):
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / header).read_text(), flags=re.S)
    out = {}
    for m in re.finditer(r"typedef struct \{([^{}]*)\} (\w+);", text):
        body, name = m.groups()
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = re.sub(r"\[[^\]]*\]", "", decl)
            fields += [re.findall(r"(\w+)\s*$", part.strip())[0] for part in decl.split(",")]
        out[name] = fields
    return out


def test_sequence_structs_match_header():
    from cuda_hlt_amd import stream
    s = _header_structs("velo_sequence.h")
    for name, st in (("velo_sequence_config", stream.SequenceConfig), ("velo_sequence_input", stream.SequenceInput),
                     ("velo_sequence_output", stream.SequenceOutput)):
        assert [f for f, _ in st._fields_] == s[name], name


def test_ghost_and_tuning_structs_match_header():
    from cuda_hlt_amd import ghost
    from cuda_hlt_amd import native as nat
    g = _header_structs("velo_ghost.h")
    for name, st in (("velo_ghost_model", ghost.ModelStruct), ("velo_track_features_args", ghost.TrackFeaturesArgs),
                     ("velo_ghost_infer_args", ghost.InferArgs), ("velo_ghost_grad_args", ghost.GradArgs),
                     ("velo_ghost_adam_args", ghost.AdamArgs)):
        assert [f for f, _ in st._fields_] == g[name], name
    h = _header_structs("velo_hlt.h")
    assert [f for f, _ in nat.Tuning._fields_] == h["velo_tuning"]
    assert [f for f, _ in nat.Constants._fields_] == h["velo_constants"]


def test_workspace_bytes_bounds():
    """velo_workspace_bytes (no GPU): the batched layout covers at least the
    bound-sized buffers, grows with events and raw bytes, and the opt-in
    Kalman and PV steps never need less (their buffers may reuse space that
    earlier steps freed)."""
    from cuda_hlt_amd.stream import workspace_bytes
    n, raw = 1000, 30_000_000
    b = workspace_bytes(n, raw)
    tracks = n * 1200
    assert b >= tracks * (54 + 26 * 16) + (raw // 2) * 24
    assert workspace_bytes(2 * n, raw) > b and workspace_bytes(n, 2 * raw) > b
    assert workspace_bytes(n, raw, velo_kalman_fit=True) >= b
    assert workspace_bytes(n, raw, pv_beamline=True) >= workspace_bytes(n, raw, velo_kalman_fit=True)
    assert b % 256 == 0
