"""The C-ABI library on the CPU: it loads, exports every declared entry point,
reports errors without exiting, and its host-side mesh and input-file API work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import euler3d_amd as E

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "euler3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(e3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(E.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in E.lib().e3d_version()


def test_mesh_api_roundtrip(tmp_path):
    m = E.Mesh.bump(6, 2, 2, 0.1, 4, (2, 1, 1))
    top = m.write(str(tmp_path), "rt")
    r = E.Mesh.read(top)
    assert r.nblocks == 2
    for b in range(2):
        x, y = m.block(b), r.block(b)
        for k in ("face2cell", "cell2face", "cell2cell", "S", "rL", "rR", "volume", "lsqw", "itf_origin_cells"):
            assert np.array_equal(x[k], y[k]), k  # %.17g text round trip is exact


def test_errors_are_reported_not_fatal(tmp_path):
    with pytest.raises(E.E3DError) as e:
        E.Mesh.read(str(tmp_path / "missing.top"))
    assert e.value.code == -3 and "missing.top" in str(e.value)
    with pytest.raises(E.E3DError) as e:
        E.Mesh.bump(0, 2, 2)
    assert e.value.code == -1
    bad = tmp_path / "bad.top"
    bad.write_text("NBLOCK= 1\n" + str(tmp_path / "nofile.su2") + "\nBlock= 0\nGhostCount= 0\nNconnexion= 0\n")
    with pytest.raises(E.E3DError):
        E.Mesh.read(str(bad))
    # an open surface: a boundary face without a marker element
    su2 = tmp_path / "open.su2"
    su2.write_text("NDIME= 3\n\nNPOIN= 4\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n\nNELEM= 1\n10 0 1 2 3\n\nNMARK= 1\n"
                   "MARKER_TAG= WALL\nMARKER_ELEMS= 1\n5 0 1 2\n")
    top = tmp_path / "open.top"
    top.write_text(f"NBLOCK= 1\n{su2}\nBlock= 0\nGhostCount= 0\nNconnexion= 0\n")
    with pytest.raises(E.E3DError) as e:
        E.Mesh.read(str(top))
    assert "unmatched" in str(e.value)


def test_read_input_file(tmp_path):
    """The reference's positional input layout (Solver/src/Interface.cpp:33-101)."""
    text = ("MESH\ntopologyfilename\n./Meshes/x.top\nINPUT\ncfl gamma angleattackdeg\n1.5 1.4 1.25\n"
            "rkstage mach cmac\n5 0.80 1.00\nnbitermax convcriterion\n100000 -15\n"
            "residual smoothing (0-no 1-yes)\n0\nfluxscheme schemeorder\nroe 2\ngradient limiter\n"
            "leastsquares venkatakrishnan\nomega k\n2.16e-16 5.00\nEXECUTABLE (0-no 1-yes)\nbuild execute\n1 1\n")
    f = tmp_path / "input"
    f.write_text(text)
    p, topo = E.read_input(str(f))
    assert topo == "./Meshes/x.top"
    assert (p.cfl, p.gamma, p.aoa_deg, p.stages, p.mach, p.cmac) == (1.5, 1.4, 1.25, 5, 0.8, 1.0)
    assert (p.max_iter, p.criterion, p.smoothing, p.flux, p.order) == (100000, -15.0, 0, E.FLUX_ROE, 2)
    assert (p.gradient, p.limiter, p.omega, p.k) == (E.GRAD_LEAST_SQUARES, E.LIM_VENKATAKRISHNAN, 2.16e-16, 5.0)
    f.write_text(text.replace("roe 2", "hllc 2"))
    with pytest.raises(E.E3DError):
        E.read_input(str(f))
    # factory fallbacks: anything but leastsquares is Green-Gauss, anything but venkatakrishnan is BJ
    f.write_text(text.replace("leastsquares venkatakrishnan", "gg bj"))
    p, _ = E.read_input(str(f))
    assert (p.gradient, p.limiter) == (E.GRAD_GREEN_GAUSS, E.LIM_BARTH_JESPERSEN)


def test_error_in_one_block_of_several_is_reported(tmp_path):
    """Blocks are built on host threads; a failure in any of them reaches the caller with its message."""
    m = E.Mesh.bump(6, 2, 2, 0.1, 4, (1, 1, 1))
    good_top = m.write(str(tmp_path), "good")
    good = [line for line in open(good_top).read().split("\n") if line.endswith(".su2")][0]
    su2 = tmp_path / "open.su2"
    su2.write_text("NDIME= 3\n\nNPOIN= 4\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n\nNELEM= 1\n10 0 1 2 3\n\nNMARK= 1\n"
                   "MARKER_TAG= WALL\nMARKER_ELEMS= 1\n5 0 1 2\n")
    blocks = [good, good, str(su2), good]
    top = tmp_path / "four.top"
    top.write_text(f"NBLOCK= {len(blocks)}\n" + "\n".join(blocks) + "\n" +
                   "".join(f"Block= {b}\nGhostCount= 0\nNconnexion= 0\n" for b in range(len(blocks))))
    with pytest.raises(E.E3DError) as e:
        E.Mesh.read(str(top))
    assert "unmatched" in str(e.value)
    top.write_text(f"NBLOCK= 2\n{good}\n{good}\n" + "".join(f"Block= {b}\nGhostCount= 0\nNconnexion= 0\n"
                                                           for b in range(2)))
    r = E.Mesh.read(str(top))
    assert r.nblocks == 2 and np.array_equal(r.block(0)["S"], r.block(1)["S"])
