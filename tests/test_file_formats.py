"""On-disk formats either side of the path, on the CPU through the C ABI:
the model dump (the reference's commented-out Vp.dat / Vs.dat,
P-S-seperating.cu:2042-2063: raw float32 interiors, x-major) and SEG-Y
revision 1 shot gathers of the seismograms.  The SEG-Y checks read the file
back with an independent parser written from the published standard (byte
positions of the binary and trace headers, big-endian IEEE samples, EBCDIC
textual header) and round-trip it through ouc_read_segy."""
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ouc_amd", "libouc.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libouc.so not built")


def test_model_files_match_reference_layout(tmp_path):
    from ouc_amd.engine import write_model_files

    rng = np.random.default_rng(1)
    vp = rng.uniform(1500, 4500, (30, 20)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    write_model_files(tmp_path / "model", vp, vs)
    for name, a in (("Vp.dat", vp), ("Vs.dat", vs)):
        raw = (tmp_path / "model" / name).read_bytes()
        assert raw == a.astype("<f4").tobytes()  # for i in x: for j in z: fwrite(float)


def test_image_files_match_reference_layout(tmp_path):
    """Per-shot Laplacian images and the stack under the reference's names
    (image_PP_lap_<nx>_<nz>_<l>.dat, All_image_PP_lap.dat, ...), each the
    interior written float by float with x outer and z inner
    (P-S-seperating.cu:2379-2400, 2446-2467).  Reading the per-shot files back
    and summing them, as the reference's stack loop does
    (P-S-seperating.cu:2403-2443), gives the stack file's values."""
    from ouc_amd.engine import write_image_files

    rng = np.random.default_rng(2)
    nx, nz = 37, 23
    shots = {l: (rng.standard_normal((nx, nz)).astype(np.float32), rng.standard_normal((nx, nz)).astype(np.float32))
             for l in (0, 3)}
    out = str(tmp_path / "image2")
    for l, (pp, ps) in shots.items():
        write_image_files(out, pp, ps, l)
    stack = [np.zeros((nx, nz), np.float32), np.zeros((nx, nz), np.float32)]
    for l in sorted(shots):
        for a, tag in enumerate(("PP", "PS")):
            raw = open(os.path.join(out, f"image_{tag}_lap_{nx}_{nz}_{l}.dat"), "rb").read()
            assert len(raw) == 4 * nx * nz
            vals = struct.unpack(f"<{nx * nz}f", raw)
            img = np.array(vals, np.float32).reshape(nx, nz)  # for i in x: for j in z
            assert np.array_equal(img, shots[l][a])
            stack[a] += img
    write_image_files(out, stack[0], stack[1], -1)
    for a, tag in enumerate(("PP", "PS")):
        raw = open(os.path.join(out, f"All_image_{tag}_lap.dat"), "rb").read()
        assert raw == stack[a].astype("<f4").tobytes()
    assert sorted(os.listdir(out)) == sorted(
        [f"image_{t}_lap_{nx}_{nz}_{l}.dat" for t in ("PP", "PS") for l in shots]
        + ["All_image_PP_lap.dat", "All_image_PS_lap.dat"])


def _be16(b, pos):  # 1-based byte position, as the standard numbers them
    return struct.unpack(">h", b[pos - 1:pos + 1])[0]


def _be32(b, pos):
    return struct.unpack(">i", b[pos - 1:pos + 3])[0]


def test_segy_gather_layout_and_round_trip(tmp_path):
    from ouc_amd.engine import read_segy, write_segy

    nx, nt, dt = 37, 150, 0.0004
    rng = np.random.default_rng(2)
    tr = rng.standard_normal((nx, nt)).astype(np.float32)
    tr[3, 5] = np.float32(1e-38)  # a subnormal survives byte for byte
    path = tmp_path / "shot7_vz.sgy"
    write_segy(path, tr, dt, shot=7, sx_m=255.0, sz_m=10.0, rec_x0_m=0.0, rec_dx_m=5.0, rec_z_m=2.5, component="VZ")
    b = path.read_bytes()
    assert len(b) == 3200 + 400 + nx * (240 + 4 * nt)
    text = b[:3200].decode("cp037")
    assert text.startswith("C 1 ") and "COMPONENT VZ" in text[80:160] and text[39 * 80:].startswith("C40")
    # binary header (file bytes 3201-3600)
    assert _be16(b, 3213) == nx
    assert _be16(b, 3217) == 400 and _be16(b, 3221) == nt
    assert _be16(b, 3225) == 5          # IEEE float samples
    assert _be16(b, 3255) == 1          # metres
    assert _be16(b, 3501) == 0x0100     # revision 1.0
    assert _be16(b, 3503) == 1          # fixed-length traces
    assert _be16(b, 3505) == 0          # no extended textual headers
    for k in (0, 1, nx - 1):
        h = b[3600 + k * (240 + 4 * nt):]
        assert _be32(h, 1) == k + 1 and _be32(h, 5) == k + 1
        assert _be32(h, 9) == 7 and _be32(h, 13) == k + 1
        assert _be16(h, 29) == 1
        assert _be16(h, 69) == -100 and _be16(h, 71) == -100
        assert _be32(h, 73) == 25500                      # source X, cm
        assert _be32(h, 81) == round(500 * k)             # group X, cm
        assert _be32(h, 37) == round(500 * k) - 25500     # offset, cm
        assert _be32(h, 41) == -250 and _be32(h, 49) == 1000
        assert _be16(h, 115) == nt and _be16(h, 117) == 400
        samples = np.frombuffer(h[240:240 + 4 * nt], ">f4")
        np.testing.assert_array_equal(samples, tr[k])
    back, dt2 = read_segy(path)
    np.testing.assert_array_equal(back, tr)
    assert dt2 == pytest.approx(dt)


def test_segy_limits(tmp_path):
    from ouc_amd import abi
    from ouc_amd.engine import OucError, read_segy, write_segy

    with pytest.raises(OucError) as e:
        write_segy(tmp_path / "long.sgy", np.zeros((2, 40000), np.float32), 0.001)
    assert e.value.status == abi.OUC_UNSUPPORTED
    (tmp_path / "bad.sgy").write_bytes(b"\0" * 100)
    with pytest.raises(OucError) as e:
        read_segy(tmp_path / "bad.sgy")
    assert e.value.status == abi.OUC_IO
