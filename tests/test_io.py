"""The io package: the xyzr sphere-pack reader against the reference's own
known answers (tests/io/test_data_read.py:8-45; the pack files under
tests/golden/sphere_packs are copies of the reference's
tests/test_data/sphere_packs data files) and the .vti writer, read back
with a parser of the VTK XML ImageData layout written here."""

import os
import re
import struct

import numpy as np
import pytest

from pmmoto_amd import initialize
from pmmoto_amd.core import BoundaryType
from pmmoto_amd.io import data_read, output

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sphere_packs")


def test_read_sphere_pack_known_answer():
    spheres, domain = data_read.read_sphere_pack_xyzr_domain(os.path.join(GOLDEN, "bcc.out"))
    np.testing.assert_array_equal(spheres, [[0.0, 0.0, 0.0, 0.25], [0.0, 0.0, 1.0, 0.25], [0.0, 1.0, 0.0, 0.25],
                                            [1.0, 0.0, 0.0, 0.25], [0.0, 1.0, 1.0, 0.25], [1.0, 0.0, 1.0, 0.25],
                                            [1.0, 1.0, 0.0, 0.25], [1.0, 1.0, 1.0, 0.25]])
    np.testing.assert_array_equal(domain, ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)))
    assert spheres.dtype == np.float64


def test_read_sphere_pack_tab_separated():
    spheres, domain = data_read.read_sphere_pack_xyzr_domain(os.path.join(GOLDEN, "10pack.out"))
    assert spheres.shape == (31, 4)  # the file holds 31 sphere lines
    np.testing.assert_allclose(domain, ((0.0, 0.394541),) * 3)
    np.testing.assert_allclose(spheres[0], [1.133208e-01, 1.990755e-01, 3.244731e-02, 8.479045e-02])


@pytest.mark.parametrize("body", ["0.5 0.5 not_a_number 0.1\n", "0.5 0.5 0.2 0.2 0.1\n"])
def test_read_sphere_pack_bad_lines(tmp_path, body):
    bad = tmp_path / "bad.xyzr"
    bad.write_text("0 1\n0 1\n0 1\n" + body)
    with pytest.raises(ValueError):
        data_read.read_sphere_pack_xyzr_domain(str(bad))


def test_read_sphere_pack_missing_file_exits(tmp_path):
    with pytest.raises(SystemExit):
        data_read.read_sphere_pack_xyzr_domain(str(tmp_path / "none.xyzr"))


def read_vti(path):
    """Parse the ImageData header and the raw appended arrays."""
    blob = open(path, "rb").read()
    head, _, rest = blob.partition(b'<AppendedData encoding="raw">\n_')
    text = head.decode()
    whole = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', text).group(1).split()]
    origin = [float(v) for v in re.search(r'Origin="([^"]*)"', text).group(1).split()]
    spacing = [float(v) for v in re.search(r'Spacing="([^"]*)"', text).group(1).split()]
    shape = tuple(whole[2 * a + 1] - whole[2 * a] for a in range(3))
    types = {"UInt8": "<u1", "Int32": "<i4", "Float32": "<f4", "Float64": "<f8", "UInt64": "<u8"}
    arrays = {}
    for name, t, off in re.findall(r'<DataArray Name="([^"]*)" NumberOfComponents="1" type="(\w+)" '
                                   r'format="appended" offset="(\d+)"/>', text):
        off = int(off)
        (nbytes,) = struct.unpack("<Q", rest[off:off + 8])
        a = np.frombuffer(rest[off + 8:off + 8 + nbytes], dtype=types[t])
        arrays[name] = a.reshape(shape, order="F")
    assert b'header_type="UInt64"' in head and b'byte_order="LittleEndian"' in head
    return whole, origin, spacing, arrays


def test_save_img_round_trip(tmp_path):
    bt = ((BoundaryType.WALL, BoundaryType.WALL), (BoundaryType.END, BoundaryType.END),
          (BoundaryType.PERIODIC, BoundaryType.PERIODIC))
    sd = initialize((6, 5, 4), box=((1.0, 4.0), (0.0, 1.0), (-2.0, 2.0)), boundary_types=bt)
    rng = np.random.default_rng(0)
    img = rng.integers(0, 2, size=sd.voxels, dtype=np.uint8)
    extra = {"d": rng.random(sd.voxels, dtype=np.float32), "lab": rng.integers(0, 9, sd.voxels).astype(np.int32)}
    path = str(tmp_path / "sub" / "img")
    output.save_img(path, sd, img, additional_img=extra)
    whole, origin, spacing, arrays = read_vti(path + ".vti")
    # pads: WALL 1, END 0, PERIODIC 1 -> start (-1, 0, -1), extent start .. start + padded voxels
    assert whole == [-1, 7, 0, 5, -1, 5]
    assert origin == [1.0, 0.0, -2.0]
    np.testing.assert_allclose(spacing, [0.5, 0.2, 1.0])
    np.testing.assert_array_equal(arrays["img"], img)
    np.testing.assert_array_equal(arrays["d"], extra["d"])
    np.testing.assert_array_equal(arrays["lab"], extra["lab"])


def test_save_img_errors(tmp_path):
    sd = initialize((10, 10, 10))
    img = np.zeros(sd.voxels)
    output.save_img(str(tmp_path / "ok"), sd, img)
    assert (tmp_path / "ok.vti").exists()
    with pytest.raises(ValueError):
        output.save_img(str(tmp_path / "a"), sd, img, additional_img={"err_img": np.zeros((5, 5, 5))})
    with pytest.raises(TypeError):
        output.save_img(str(tmp_path / "b"), sd, img, additional_img={"img2": 4.0})
    with pytest.raises(NotImplementedError):
        output.save_img(str(tmp_path / "c"), {0: sd}, {0: img})
