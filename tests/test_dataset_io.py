"""Dataset-format IO (the reference's OpenImageIO reads and PNG writes,
opencl/bmfr.cpp:145-172, 519-547, and its compiled-in camera_matrices.h,
opencl/bmfr.cpp:43-47, 226-227, 440-444).

The EXR files here are produced by an encoder written in this test from the
OpenEXR file layout (header attributes, offset table, scanline chunks, the
ZIP/RLE byte reorder and predictor) -- independent of the library's reader,
which only decodes.  PNGs are decoded with Pillow.  The GPU test runs the
reference's frame loop over a dataset directory written from the procedural
scene and checks it against the same frames driven directly."""
import os
import struct
import zlib

import numpy as np
import pytest

import bmfr_amd
from bmfr_amd import dataset

HALF, FLOAT, UINT = 1, 2, 0


def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def _split_and_predict(raw):
    """The compressor side of OpenEXR's ZIP/RLE pre-pass: even bytes to the
    first half and odd bytes to the second, then each byte replaced by its
    difference to the previous one plus 128."""
    t = raw[0::2] + raw[1::2]
    out = bytearray(t)
    for i in range(1, len(t)):
        out[i] = (t[i] - t[i - 1] + 128) & 0xFF
    return bytes(out)


def _rle(data):
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i
        while j < n and j - i < 127 and data[j] == data[i]:
            j += 1
        if j - i >= 3:
            out += bytes([j - i - 1, data[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 127 and not (j + 2 < n and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        out += bytes([(256 - (j - i)) & 0xFF]) + data[i:j]
        i = j
    return bytes(out)


def encode_exr(channels, compression, xmin=0, ymin=0):
    """channels: {name: (H, W) array}; pixel type from the dtype (float16
    HALF, float32 FLOAT, uint32 UINT)."""
    names = sorted(channels)
    H, W = channels[names[0]].shape
    kinds = {np.dtype(np.float16): HALF, np.dtype(np.float32): FLOAT, np.dtype(np.uint32): UINT}
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", kinds[channels[n].dtype], 0, 1, 1) for n in names)
    win = struct.pack("<4i", xmin, ymin, xmin + W - 1, ymin + H - 1)
    header = (struct.pack("<II", 20000630, 2) + _attr("channels", "chlist", chl + b"\0")
              + _attr("compression", "compression", bytes([compression]))
              + _attr("dataWindow", "box2i", win) + _attr("displayWindow", "box2i", win)
              + _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1))
              + _attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    lpc = 16 if compression == 3 else 1
    chunks = []
    for y0 in range(0, H, lpc):
        raw = b"".join(channels[n][y].astype(channels[n].dtype.newbyteorder("<")).tobytes()
                       for y in range(y0, min(H, y0 + lpc)) for n in names)
        if compression == 0:
            data = raw
        elif compression == 1:
            data = _rle(_split_and_predict(raw))
        else:
            data = zlib.compress(_split_and_predict(raw))
        if len(data) >= len(raw):
            data = raw
        chunks.append(struct.pack("<ii", ymin + y0, len(data)) + data)
    table_end = len(header) + 8 * len(chunks)
    offs, pos = [], table_end
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    return header + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks)


def _image(h, w, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((h, w, 3)).astype(np.float32) * 4
    a[0, :5] = 0.0                        # runs for RLE
    a[1, :, 0] = 1.25
    return a


@pytest.mark.parametrize("comp", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["half", "float"])
def test_exr_reader_against_independent_encoder(tmp_path, comp, kind):
    h, w = 37, 29   # ZIP chunks of 16 lines with a short last chunk
    img = _image(h, w, comp * 7 + len(kind))
    dt = np.float16 if kind == "half" else np.float32
    chans = {c: img[:, :, i].astype(dt) for i, c in enumerate("RGB")}
    chans["A"] = np.ones((h, w), dt)        # an extra channel, ignored
    p = tmp_path / f"t{comp}{kind}.exr"
    p.write_bytes(encode_exr(chans, comp, xmin=3, ymin=-2))
    assert dataset.exr_info(p) == (w, h, 4)
    got = dataset.read_exr_rgb(p)
    want = np.stack([chans[c].astype(np.float32) for c in "RGB"], axis=-1)
    assert np.array_equal(got, want)


def test_exr_reader_half_specials_uint_and_luminance(tmp_path):
    h, w = 3, 8
    bits = np.array([0x0000, 0x8000, 0x0001, 0x03FF, 0x0400, 0x3C00, 0x7BFF, 0x7C00,
                     0xFC00, 0x7E00, 0xC000, 0x3555, 0x0200, 0x8001, 0x4248, 0x5640] * 2, np.uint16)[: h * w]
    y = bits.view(np.float16).reshape(h, w)
    p = tmp_path / "y.exr"
    p.write_bytes(encode_exr({"Y": y}, 2))
    got = dataset.read_exr_rgb(p)
    want = y.astype(np.float32)
    for c in range(3):
        assert np.array_equal(got[:, :, c], want, equal_nan=True)
    u = (np.arange(h * w, dtype=np.uint32) * 1234567).reshape(h, w)
    p2 = tmp_path / "u.exr"
    p2.write_bytes(encode_exr({"R": u, "G": u, "B": u}, 0))
    assert np.array_equal(dataset.read_exr_rgb(p2)[:, :, 1], u.astype(np.float32))


@pytest.mark.parametrize("comp", [dataset.EXR_NONE, dataset.EXR_ZIP])
def test_exr_write_read_round_trip(tmp_path, comp):
    img = _image(41, 23, 5)
    img[2, 3] = [np.inf, -np.inf, np.nan]
    p = tmp_path / "rt.exr"
    dataset.write_exr_rgb(p, img, comp)
    assert dataset.exr_info(p) == (23, 41, 3)
    assert np.array_equal(dataset.read_exr_rgb(p), img, equal_nan=True)


def test_exr_refusals(tmp_path):
    bad = tmp_path / "bad.exr"
    bad.write_bytes(b"not an exr file at all")
    with pytest.raises(bmfr_amd.BmfrError):
        dataset.exr_info(bad)
    img = {c: np.zeros((4, 4), np.float16) for c in "RGB"}
    piz = tmp_path / "piz.exr"
    data = bytearray(encode_exr(img, 0))
    i = data.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    data[i] = 4   # PIZ
    piz.write_bytes(bytes(data))
    with pytest.raises(bmfr_amd.BmfrError, match="compression 4"):
        dataset.read_exr_rgb(piz)
    tiled = tmp_path / "tiled.exr"
    d2 = bytearray(encode_exr(img, 0))
    d2[4:8] = struct.pack("<I", 2 | 0x200)
    tiled.write_bytes(bytes(d2))
    with pytest.raises(bmfr_amd.BmfrError, match="scanline"):
        dataset.exr_info(tiled)
    ok = tmp_path / "ok.exr"
    ok.write_bytes(encode_exr(img, 0))
    with pytest.raises(bmfr_amd.BmfrError, match="expected"):
        dataset.read_exr_rgb(ok, 5, 4)


@pytest.mark.parametrize("comp", [0, 1, 3])
def test_exr_reader_survives_corrupted_files(tmp_path, comp):
    """Seeded byte mutations, truncations and oversized fields of a valid
    file: the reader either reads an image of the header's size or returns
    a status; it never reads or writes out of bounds (run under the host
    AddressSanitizer build as well) and never crashes."""
    h, w = 20, 11
    img = _image(h, w, 3 + comp)
    chans = {c: img[:, :, i].astype(np.float16) for i, c in enumerate("RGB")}
    good = encode_exr(chans, comp)
    rng = np.random.default_rng(100 + comp)
    p = tmp_path / "m.exr"
    outcomes = {"read": 0, "refused": 0}
    for trial in range(150):
        d = bytearray(good)
        kind = trial % 3
        if kind == 0:   # a few random bytes anywhere
            for _ in range(int(rng.integers(1, 6))):
                d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
        elif kind == 1:   # truncated
            d = d[:int(rng.integers(8, len(d)))]
        else:   # a 32-bit field (size, offset, line number, window) set to an extreme
            at = int(rng.integers(8, len(d) - 4))
            d[at:at + 4] = struct.pack("<i", int(rng.choice([-1, 2**31 - 1, -2**31, 1 << 24, 0])))
        p.write_bytes(bytes(d))
        try:
            ww, hh, _ = dataset.exr_info(p)
            out = dataset.read_exr_rgb(p, ww, hh)
            assert out.shape == (hh, ww, 3)
            outcomes["read"] += 1
        except bmfr_amd.BmfrError:
            outcomes["refused"] += 1
    assert outcomes["refused"] > 0 and sum(outcomes.values()) == 150


@pytest.mark.parametrize("comp,lines_per_chunk", [(0, 1), (3, 16)])
@pytest.mark.parametrize("offset", [2**64 - 1, 2**64 - 4, 2**63, None])
def test_exr_chunk_offset_out_of_file_is_refused(tmp_path, comp, lines_per_chunk, offset):
    """A chunk offset at or past the end of the file, including values whose
    sum with a field size wraps a 64-bit index, is refused with a status."""
    h, w = 20, 11
    img = _image(h, w, 7)
    good = bytearray(encode_exr({c: img[:, :, i].astype(np.float16) for i, c in enumerate("RGB")}, comp))
    chunks = -(-h // lines_per_chunk)
    # the offset table follows the header; its first entry points just past it
    table = next(pos for pos in range(8, len(good) - 8)
                 if struct.unpack_from("<Q", good, pos)[0] == pos + 8 * chunks)
    bad = bytearray(good)
    struct.pack_into("<Q", bad, table + 8 * (chunks - 1), len(good) if offset is None else offset)
    p = tmp_path / "m.exr"
    p.write_bytes(bytes(bad))
    with pytest.raises(bmfr_amd.BmfrError) as e:
        dataset.read_exr_rgb(p, w, h)
    assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT
    p.write_bytes(bytes(good))
    assert dataset.read_exr_rgb(p, w, h).shape == (h, w, 3)


def test_png_writer_decodes_with_pillow(tmp_path):
    from PIL import Image
    img = _image(13, 17, 9) * 0.3 + 0.5
    img[0, 0] = [np.nan, -1.0, 2.0]
    p = tmp_path / "o.png"
    dataset.write_png_rgb(p, img)
    got = np.asarray(Image.open(p).convert("RGB"))
    want = (np.clip(np.nan_to_num(img, nan=0.0), 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert got.shape == (13, 17, 3)
    assert np.array_equal(got, want)


CAMERA_H = """
// camera data exported for 3 frames
#pragma once
#define position_limit_squared 0.003600f
const float normal_limit_squared = 1e-1f;   /* as printed with 6 digits */
static const float camera_matrices[3][4][4] = {
  { {1.0f, 0, 0, 0}, {0, 1.f, 0, 0}, {0, 0, -1.0002f, -1}, {0.5f, -2.5e-1f, 3.0f, 4.0f} },
  { {2.0f, 0, 0, 0}, {0, 2.f, 0, 0}, {0, 0, -1.0002f, -1}, {0.5f, -2.5E-1F, 3.0f, 5.0f} },
  { {3.0f, 0, 0, 0}, {0, 3.f, 0, 0}, {0, 0, -1.0002f, -1}, {.5f, -2.5e-1f, 3.0f, 6.0f} },
};
const float pixel_offsets[3][2] = { {0.5f, 0.5f}, {0.25f, -0.125f}, {1e-3f, 0.75f} };
"""


def test_camera_header_parser(tmp_path):
    p = tmp_path / "camera_matrices.h"
    p.write_text(CAMERA_H)
    m, o, pl, nl = dataset.load_camera(p, 3)
    assert m.shape == (3, 16)
    np.testing.assert_array_equal(m[:, 0], [1, 2, 3])
    np.testing.assert_array_equal(m[1, 10:], np.float32([-1.0002, -1, 0.5, -0.25, 3, 5]))
    np.testing.assert_array_equal(o, np.float32([[0.5, 0.5], [0.25, -0.125], [1e-3, 0.75]]))
    assert pl == np.float32(0.0036) and nl == np.float32(0.1)
    with pytest.raises(bmfr_amd.BmfrError, match="values"):
        dataset.load_camera(p, 4)
    q = tmp_path / "partial.h"
    q.write_text("float camera_matrices[1][16] = {" + ",".join(["0"] * 16) + "};")
    with pytest.raises(bmfr_amd.BmfrError, match="pixel_offsets"):
        dataset.load_camera(q, 1)


def _write_dataset(d, sc, w, h, frames):
    lines = ["// written from the procedural scene", "#define position_limit_squared 0.01",
             "#define normal_limit_squared 0.1", f"float camera_matrices[{frames}][4][4] = {{"]
    offs = []
    for f in range(frames):
        g = sc.render_host(f)
        for key, stem in (("albedo", dataset.ALBEDO), ("normals", dataset.NORMAL),
                          ("positions", dataset.POSITION), ("noisy", dataset.NOISY)):
            dataset.write_exr_rgb(os.path.join(d, f"{stem}{f}.exr"), g[key].reshape(h, w, 3))
        # camera_matrices[f] is frame f's own matrix; the frame loop binds
        # camera_matrices[max(f - 1, 0)] (bmfr.cpp:440-442)
        vp, off = sc.camera(f)
        lines.append("{" + ",".join(repr(float(v)) + "f" for v in vp) + "},")
        offs.append(off)
    lines.append("};")
    lines.append(f"float pixel_offsets[{frames}][2] = {{" +
                 ",".join("{%r, %r}" % (float(o[0]), float(o[1])) for o in offs) + "};")
    open(os.path.join(d, "camera_matrices.h"), "w").write("\n".join(lines) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("features", [None, ("1.f,normal.y,", "world_position.x*world_position.y,world_position.z")])
def test_dataset_run_equals_direct_frames(gpu, tmp_path, features):
    w, h, frames = 96, 64, 4
    sc = bmfr_amd.Scene(0xB3F40000, 8, 1e-3, w, h)
    _write_dataset(str(tmp_path), sc, w, h, frames)
    out = tmp_path / "outputs"
    res = dataset.run(str(tmp_path), frames, str(out), features=features)
    p = bmfr_amd.default_params(w, h)
    if features is not None:
        bmfr_amd.parse_feature_buffers(p, *features)
        assert p.feature_count == 4
    st = bmfr_amd.Stream(p)
    try:
        for f in range(frames):
            g = sc.render_host(f)
            st.upload_frame(g["albedo"], g["normals"], g["positions"], g["noisy"])
            vp, off = sc.frame_args(f)
            st.run_frame(f, vp, off)
            assert np.array_equal(st.download_result().reshape(h, w, 3), res[f], equal_nan=True), f
    finally:
        st.close()
    assert sorted(os.listdir(out)) == [f"output{f}.png" for f in range(frames)]
