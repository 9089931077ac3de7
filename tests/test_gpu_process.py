"""Compressor::process (the InputOptions pipeline, Context.cpp:217-344) of the
nvtt.h front end, driven from C++ (tests/cpp/nvtt_api_test.cpp "process"),
against the oracle running the reference's steps in the reference's order:
setImage -> toLinear(inputGamma) -> buildNextMipmap in linear space (or the
supplied level, made linear) -> [normal map: expand, normalise, pack] ->
toGamma(outputGamma) copy -> compress, face by face, level by level.  The DDS
header is checked field by field against the DirectDraw Surface layout and
the writer's fields (DirectDrawSurface.cpp:574-813, Context.cpp:604-868)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import learned, synth

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(N.LIB_DIR, "nvtt_api_test")
C_DRIVER = os.path.join(N.LIB_DIR, "nvtt_wrapper_test")
DDS, DDS10, KTX = 0, 1, 2
UNSUPPORTED_FEATURE, UNSUPPORTED_OUTPUT = 2, 6


def fourcc(s):
    return struct.unpack("<I", s.encode("ascii"))[0]


def unit(g):
    return abs(g - 1.0) <= 1e-4 * max(1.0, abs(g))


def mip_count(w, h, mipmaps, max_level):
    n = len(N.mip_sizes(w, h)) if mipmaps else 1
    return min(n, max_level) if (mipmaps and max_level > 0) else n


def dds_header(fmt, w, h, mips, cube, normal, container, version=7):
    flags = 0x1 | 0x1000 | 0x4 | 0x2                      # CAPS | PIXELFORMAT | WIDTH | HEIGHT
    caps1, caps2 = 0x1000, 0                               # TEXTURE
    if cube:
        caps1 |= 0x8
        caps2 = 0x200 | 0xFC00                             # CUBEMAP | ALL_FACES
    if mips > 1:
        flags |= 0x20000                                   # MIPMAPCOUNT
        caps1 |= 0x8 | 0x400000                            # COMPLEX | MIPMAP
        count = mips
    else:
        count = 1
        caps1 = 0x1000 if caps2 == 0 else 0x1000 | 0x8
    reserved = [0] * 11
    reserved[7], reserved[8] = fourcc("UVER"), version
    reserved[9], reserved[10] = fourcc("NVTT"), (2 << 16) | (1 << 8) | 2
    pitch, pf_flags, pf_fourcc, pf_bits = 0, 0x4, 0, 0     # FOURCC
    if container == DDS10:
        pf_fourcc = fourcc("DX10")
        if normal and fmt in (N.FORMAT_BC1, N.FORMAT_BC5, N.FORMAT_BC3N):
            pf_flags |= 0x80000000
    else:
        flags |= 0x80000                                   # LINEARSIZE
        pitch = N.compressed_size(fmt, w, h)
        pf_fourcc = fourcc({N.FORMAT_BC1: "DXT1", N.FORMAT_BC3: "DXT5", N.FORMAT_BC3N: "DXT5", N.FORMAT_BC4: "ATI1",
                            N.FORMAT_BC5: "ATI2"}[fmt])
        if normal and fmt == N.FORMAT_BC1:
            pf_flags |= 0x80000000
        if normal and fmt == N.FORMAT_BC5:
            pf_flags |= 0x80000000
            pf_bits = fourcc("A2XY")
        if normal and fmt == N.FORMAT_BC3N:
            pf_flags |= 0x80000000
            pf_bits = fourcc("A2D5")
    words = [fourcc("DDS "), 124, flags, h, w, pitch, 0, count] + reserved
    words += [32, pf_flags, pf_fourcc, pf_bits, 0, 0, 0, 0, caps1, caps2, 0, 0, 0]
    if container == DDS10:
        words += [{N.FORMAT_BC1: 71, N.FORMAT_BC3: 77, N.FORMAT_BC3N: 77, N.FORMAT_BC4: 80, N.FORMAT_BC5: 83}[fmt], 3,
                  4 if cube else 0, 1, 0]
    return struct.pack("<%dI" % len(words), *words)


def run_process(tmp_path, faces, fmt, filt, cube=False, alpha_mode=0, wrap=2, normal=False, gammas=(2.2, 2.2),
                container=DDS, header=True, mipmaps=True, max_level=-1, normalize=True, driver=DRIVER, extra=()):
    """faces: per face, the supplied levels (level 0 first) as HxWx4 RGBA8.
    driver: the C++ front-end driver, or C_DRIVER for the C interface (which
    has no user-version setter, so its DDS headers carry version 0); extra:
    C_DRIVER's quality, colour-dithering and binary-alpha arguments."""
    h, w = faces[0][0].shape[:2]
    src, dst = tmp_path / "in.raw", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(struct.pack("<iiii", w, h, len(faces), len(faces[0])))
        for lv in faces:
            for img in lv:
                f.write(synth.to_bgra(img).tobytes())
    args = [driver, "process", str(src), str(dst), str(fmt), str(filt), str(int(cube)), str(alpha_mode), str(wrap),
            str(int(normal)), repr(gammas[0]), repr(gammas[1]), str(container), str(int(header)), str(int(mipmaps)),
            str(max_level), str(int(normalize))] + [str(x) for x in extra]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = dst.read_bytes()
    (nb,) = struct.unpack_from("<i", data, 0)
    stream = data[4:4 + nb]
    off = 4 + nb
    (ne,) = struct.unpack_from("<i", data, off)
    events = np.frombuffer(data, np.int32, 7 * ne, off + 4).reshape(ne, 7)
    off += 4 + 28 * ne
    (nerr,) = struct.unpack_from("<i", data, off)
    errors = list(np.frombuffer(data, np.int32, nerr, off + 4))
    off += 4 + 4 * nerr
    ok, est = struct.unpack_from("<ii", data, off)
    return stream, events, errors, bool(ok), est


def oracle_process(oracle, faces, fmt, filt, alpha_mode=0, wrap=2, normal=False, gammas=(2.2, 2.2), levels=1,
                   normalize=True, ktx=False, quality=None):
    """The reference's steps; ktx: its KTX branch (Context.cpp:345-474), which
    passes the Kaiser parameters as (stretch, alpha) and only normalises
    normal-map levels.  Every supplied level is used in both cases here (the
    tests supply leading levels only)."""
    width, params = N.DEFAULT_FILTER[filt]
    if ktx and filt == N.FILTER_KAISER:
        params = (params[1], params[0])
    out = []
    for lv in faces:
        h, w = lv[0].shape[:2]
        lin = None
        for m, (lw, lh) in enumerate(N.mip_sizes(w, h)[:levels]):
            if m < len(lv):
                x = synth.to_planar(lv[m]).reshape(-1)
                lin = x if (normal or unit(gammas[0])) else oracle.planar_to_linear(x, gammas[0])
            else:
                pw, ph = N.mip_sizes(w, h)[m - 1]
                if alpha_mode == N.ALPHA_TRANSPARENCY:
                    lin = oracle.downsample(filt, width, params, wrap, lin, pw, ph)
                else:
                    lin = oracle.build_next_mipmap(filt, width, params, wrap, lin, pw, ph)
            if m > 0 and normal and normalize:
                lin = oracle.normalize_vectors(lin) if ktx else oracle.normalize_normals(lin)
            tmp = lin if (normal or unit(gammas[1])) else oracle.planar_to_gamma(lin, gammas[1])
            if quality is None:
                out.append((lw, lh, oracle.compress(fmt, tmp, lw, lh, alpha_mode=alpha_mode)))
            else:
                out.append((lw, lh, oracle.compress_q(fmt, quality, tmp, lw, lh, alpha_mode=alpha_mode)))
    return out


def check(stream, events, errors, ok, est, expect_header, expect, faces_n, fmt):
    assert ok and errors == []
    hb = len(expect_header)
    assert stream[:hb] == expect_header
    off = hb
    ev = 0
    if hb:
        assert list(events[0][:2]) == [2, hb]
        ev = 1
    levels = len(expect) // faces_n
    for i, (lw, lh, blocks) in enumerate(expect):
        face, mip = divmod(i, levels)
        got = np.frombuffer(stream, np.uint8, blocks.size, off)
        assert np.array_equal(got, blocks), (face, mip, lw, lh)
        off += blocks.size
        assert list(events[ev]) == [1, blocks.size, lw, lh, 1, face, mip]
        assert list(events[ev + 1][:2]) == [2, blocks.size]
        assert events[ev + 2][0] == 3
        ev += 3
    assert off == len(stream) and ev == len(events)
    assert est == sum(b.size for _, _, b in expect)


CASES = [
    # id, format, filter, size, kwargs
    ("bc1-box-default", N.FORMAT_BC1, N.FILTER_BOX, (128, 96), {}),
    ("bc3-kaiser-alpha", N.FORMAT_BC3, N.FILTER_KAISER, (70, 33), dict(alpha_mode=N.ALPHA_TRANSPARENCY, wrap=N.WRAP_CLAMP)),
    ("bc1-triangle-dds10-linear", N.FORMAT_BC1, N.FILTER_TRIANGLE, (64, 40), dict(gammas=(1.0, 1.0), container=DDS10)),
    ("bc5-normal-kaiser", N.FORMAT_BC5, N.FILTER_KAISER, (64, 64), dict(normal=True)),
    ("bc5-normal-raw-mips", N.FORMAT_BC5, N.FILTER_BOX, (32, 48), dict(normal=True, normalize=False, container=DDS10)),
    ("bc4-single-level", N.FORMAT_BC4, N.FILTER_BOX, (96, 64), dict(mipmaps=False)),
    ("bc3n-normal-kaiser", N.FORMAT_BC3N, N.FILTER_KAISER, (64, 48), dict(normal=True)),
    ("bc3n-normal-dds10", N.FORMAT_BC3N, N.FILTER_TRIANGLE, (40, 40), dict(normal=True, container=DDS10)),
    ("bc1-max-level-no-header", N.FORMAT_BC1, N.FILTER_KAISER, (100, 60), dict(max_level=3, header=False)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_process_matches_oracle(tmp_path, oracle, case):
    _, fmt, filt, (w, h), kw = case
    if fmt in (N.FORMAT_BC5, N.FORMAT_BC3N):
        img = synth.normal_map_rgba(w, h, seed=8)
    elif fmt == N.FORMAT_BC4:
        img = synth.single_channel_rgba(w, h, seed=9)
    else:
        img = synth.noise_rect_rgba(w, h, seed=10, transparent=True)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], fmt, filt, **kw)
    levels = mip_count(w, h, kw.get("mipmaps", True), kw.get("max_level", -1))
    okw = {k: v for k, v in kw.items() if k in ("alpha_mode", "wrap", "normal", "gammas", "normalize")}
    expect = oracle_process(oracle, [[img]], fmt, filt, levels=levels, **okw)
    hdr = dds_header(fmt, w, h, levels, False, kw.get("normal", False), kw.get("container", DDS)) \
        if kw.get("header", True) else b""
    check(stream, events, errors, ok, est, hdr, expect, 1, fmt)


def test_process_cube_map(tmp_path, oracle):
    faces = [[synth.noise_rect_rgba(32, 32, seed=20 + f)] for f in range(6)]
    stream, events, errors, ok, est = run_process(tmp_path, faces, N.FORMAT_BC1, N.FILTER_KAISER, cube=True)
    expect = oracle_process(oracle, faces, N.FORMAT_BC1, N.FILTER_KAISER, levels=6)
    check(stream, events, errors, ok, est, dds_header(N.FORMAT_BC1, 32, 32, 6, True, False, DDS), expect, 6,
          N.FORMAT_BC1)


def test_process_supplied_levels(tmp_path, oracle):
    """Levels 0 and 1 supplied (each made linear), the rest built from level 1."""
    l0 = synth.noise_rect_rgba(64, 32, seed=30)
    l1 = synth.noise_rect_rgba(32, 16, seed=31)
    stream, events, errors, ok, est = run_process(tmp_path, [[l0, l1]], N.FORMAT_BC3, N.FILTER_TRIANGLE)
    expect = oracle_process(oracle, [[l0, l1]], N.FORMAT_BC3, N.FILTER_TRIANGLE, levels=7)
    check(stream, events, errors, ok, est, dds_header(N.FORMAT_BC3, 64, 32, 7, False, False, DDS), expect, 1,
          N.FORMAT_BC3)


def test_process_other_gamma(tmp_path, oracle):
    """Gamma 1.8: the 8-bit input goes through a host-built table (exact); the
    toGamma copies are correctly rounded on the device where the reference's
    C library powf is off by one ulp for about 5 in 10^4 values, which can move
    a block: at least 99% of the blocks identical, decoded PSNR within 0.01 dB."""
    img = synth.noise_rect_rgba(128, 128, seed=40)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], N.FORMAT_BC1, N.FILTER_BOX, gammas=(1.8, 1.8))
    expect = oracle_process(oracle, [[img]], N.FORMAT_BC1, N.FILTER_BOX, gammas=(1.8, 1.8), levels=8)
    assert ok and errors == []
    got = np.frombuffer(stream, np.uint8)[128:]
    exp = np.concatenate([b for _, _, b in expect])
    same = np.mean(np.all(got.reshape(-1, 8) == exp.reshape(-1, 8), axis=1))
    assert same >= 0.99, same
    planar = synth.to_planar(img).reshape(-1)
    n0 = N.compressed_size(N.FORMAT_BC1, 128, 128)
    p_got = learned.bc1_psnr(got[:n0], planar, 128, 128)
    p_exp = learned.bc1_psnr(exp[:n0], planar, 128, 128)
    assert abs(p_got - p_exp) <= 0.01, (p_got, p_exp)


def test_process_unsupported_output_format(tmp_path):
    img = synth.noise_rect_rgba(16, 16, seed=1)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], 11, N.FILTER_BOX)  # Format_BC7
    assert not ok and errors == [UNSUPPORTED_OUTPUT] and stream == b"" and len(events) == 0


def test_process_learned_bc1(tmp_path, oracle, gpu):
    """CompressionOptions::setLearnedBC1 (an extension of the front end):
    process() compresses the levels by the learned form with the per-level
    settings of learned.level_settings (shipped at level 0, refined at level
    1, the plain cluster fit from 4096 blocks down).  The stream equals the
    multi-image learned entry point's output for the same network on the
    pipeline's levels (the oracle's, which the other tests pin to the
    device's), and level 0 differs from the plain fit (the learned path ran)."""
    w = h = 512
    img = synth.noise_rect_rgba(w, h, seed=21)
    rng = np.random.default_rng(4)
    with N.Mlp(gpu, seed=5) as net:
        params = net.params()
    # A network whose output layer moves the candidates (the seeded start has
    # a zero output layer, which is exactly the box-diagonal method).
    w3 = 3072 + 64 + 4096 + 64
    params[w3:w3 + 25 * 64] = rng.normal(0.0, 0.05, 25 * 64).astype(np.float32)
    pfile = tmp_path / "params.bin"
    params.astype(np.float32).tofile(pfile)
    levels = len(N.mip_sizes(w, h))
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], N.FORMAT_BC1, N.FILTER_BOX, header=False,
                                                  extra=(str(pfile),))
    assert ok and errors == []
    # The pipeline's levels: to linear, Box mips, back to gamma 2.2.
    lin = oracle.planar_to_linear(synth.to_planar(img).reshape(-1), 2.2)
    sizes = N.mip_sizes(w, h)
    gam = []
    for m, (lw, lh) in enumerate(sizes):
        if m:
            lin = oracle.build_next_mipmap(N.FILTER_BOX, 0.5, None, N.WRAP_MIRROR, lin, *sizes[m - 1])
        gam.append(oracle.planar_to_gamma(lin, 2.2))
    d_in = [gpu.alloc(x.nbytes).upload(x) for x in gam]
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, a, b)) for a, b in sizes]
    with N.Mlp(gpu, params) as net:
        net.compress_bc1_images_device(d_in, sizes, d_out, levels=list(range(levels)))
    expect = [d.download(np.uint8) for d in d_out]
    assert [learned.level_settings(a, b, m) is None for m, (a, b) in enumerate(sizes)][:3] == [False, False, True]
    off = 0
    for m, blocks in enumerate(expect):
        got = np.frombuffer(stream, np.uint8, blocks.size, off)
        assert np.array_equal(got, blocks), m
        off += blocks.size
    assert off == len(stream)
    assert not np.array_equal(expect[0], oracle.compress(N.FORMAT_BC1, gam[0], w, h))
    for d in d_in + d_out:
        d.free()


def ktx_header(fmt, w, h, mips, faces):
    """KTX 1.1 header (the Khronos layout; KtxFile.h:113-131, Context.cpp:870-1030)."""
    internal, base = {N.FORMAT_BC1: (0x83F0, 0x1907), N.FORMAT_BC3: (0x83F3, 0x1908), N.FORMAT_BC3N: (0x83F3, 0x1908),
                      N.FORMAT_BC4: (0x8DBB, 0x1903), N.FORMAT_BC5: (0x8DBD, 0x8227)}[fmt]
    ident = bytes([0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31, 0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A])
    return ident + struct.pack("<13I", 0x04030201, 0, 1, 0, internal, base, w, h, 0, 0, faces, mips, 0)


KTX_CASES = [
    ("bc1-kaiser-2d", N.FORMAT_BC1, N.FILTER_KAISER, False, False, (64, 48)),
    ("bc3-box-cube", N.FORMAT_BC3, N.FILTER_BOX, True, False, (16, 16)),
    ("bc5-normal-triangle", N.FORMAT_BC5, N.FILTER_TRIANGLE, False, True, (32, 32)),
    ("bc4-kaiser", N.FORMAT_BC4, N.FILTER_KAISER, False, False, (40, 24)),
    ("bc3n-normal-box", N.FORMAT_BC3N, N.FILTER_BOX, False, True, (24, 40)),
]


@pytest.mark.parametrize("case", KTX_CASES, ids=[c[0] for c in KTX_CASES])
def test_process_ktx(tmp_path, oracle, case):
    """The KTX branch: header, then per level a uint32 byte count of all faces,
    that level of every face, padding after levels past the first."""
    _, fmt, filt, cube, normal, (w, h) = case
    nf = 6 if cube else 1
    if fmt in (N.FORMAT_BC5, N.FORMAT_BC3N):
        faces = [[synth.normal_map_rgba(w, h, seed=50 + f)] for f in range(nf)]
    elif fmt == N.FORMAT_BC4:
        faces = [[synth.single_channel_rgba(w, h, seed=50 + f)] for f in range(nf)]
    else:
        faces = [[synth.noise_rect_rgba(w, h, seed=50 + f, transparent=True)] for f in range(nf)]
    stream, events, errors, ok, est = run_process(tmp_path, faces, fmt, filt, cube=cube, normal=normal, container=KTX)
    assert ok and errors == []
    levels = len(N.mip_sizes(w, h))
    expect = oracle_process(oracle, faces, fmt, filt, normal=normal, levels=levels, ktx=True)
    exp = ktx_header(fmt, w, h, levels, nf)
    ev = [(2, 64)]
    for m, (lw, lh) in enumerate(N.mip_sizes(w, h)):
        size = N.compressed_size(fmt, lw, lh)
        exp += struct.pack("<I", size * nf)
        ev.append((2, 4))
        for f in range(nf):
            exp += expect[f * levels + m][2].tobytes()
            ev += [(1, size), (2, size), (3, 0)]
        if m > 0 and (3 - (size * nf + 3) % 4):
            exp += bytes(3 - (size * nf + 3) % 4)
            ev.append((2, 3 - (size * nf + 3) % 4))
    assert stream == exp
    assert [tuple(e[:2]) if e[0] != 3 else (3, 0) for e in events] == ev
    assert est == nf * sum(N.compressed_size(fmt, a, b) for a, b in N.mip_sizes(w, h))


C_CASES = [c for c in CASES if c[0] in ("bc1-box-default", "bc3-kaiser-alpha", "bc5-normal-kaiser", "bc4-single-level")]


@pytest.mark.parametrize("case", C_CASES, ids=[c[0] for c in C_CASES])
def test_c_interface_process(tmp_path, oracle, case):
    """nvttCompress (nvtt_wrapper.cpp:313-316) from a plain C caller: the same
    stream, handler events and size estimate as Compressor::process."""
    _, fmt, filt, (w, h), kw = case
    if fmt == N.FORMAT_BC5:
        img = synth.normal_map_rgba(w, h, seed=8)
    elif fmt == N.FORMAT_BC4:
        img = synth.single_channel_rgba(w, h, seed=9)
    else:
        img = synth.noise_rect_rgba(w, h, seed=10, transparent=True)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], fmt, filt, driver=C_DRIVER, **kw)
    levels = mip_count(w, h, kw.get("mipmaps", True), kw.get("max_level", -1))
    okw = {k: v for k, v in kw.items() if k in ("alpha_mode", "wrap", "normal", "gammas", "normalize")}
    expect = oracle_process(oracle, [[img]], fmt, filt, levels=levels, **okw)
    hdr = dds_header(fmt, w, h, levels, False, kw.get("normal", False), kw.get("container", DDS), version=0)
    check(stream, events, errors, ok, est, hdr, expect, 1, fmt)
    cpp = run_process(tmp_path, [[img]], fmt, filt, **kw)
    assert cpp[0][len(hdr):] == stream[len(hdr):]
    assert np.array_equal(cpp[1], events) and cpp[4] == est


def test_c_interface_ktx_production(tmp_path, oracle):
    """The C interface with a quality set: BC1 Production (icbc Level9) into
    KTX, every level against the oracle at that quality."""
    w, h = 48, 32
    img = synth.noise_rect_rgba(w, h, seed=60)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], N.FORMAT_BC1, N.FILTER_BOX, driver=C_DRIVER,
                                                  extra=(2, 0, 0), container=KTX)
    assert ok and errors == []
    sizes = N.mip_sizes(w, h)
    expect = oracle_process(oracle, [[img]], N.FORMAT_BC1, N.FILTER_BOX, levels=len(sizes), ktx=True, quality=2)
    exp = ktx_header(N.FORMAT_BC1, w, h, len(sizes), 1)
    for m, (lw, lh, blocks) in enumerate(expect):
        exp += struct.pack("<I", blocks.size) + blocks.tobytes()
        if m > 0:
            exp += bytes(3 - (blocks.size + 3) % 4)
    assert stream == exp
    assert est == sum(b.size for _, _, b in expect)


@pytest.mark.parametrize("fmt,dither,binary", [(N.FORMAT_BC1, 1, 0), (N.FORMAT_BC3, 1, 0), (N.FORMAT_BC4, 0, 1)])
def test_c_interface_quantize_unsupported(tmp_path, fmt, dither, binary):
    """Colour dithering (BC1-BC3) and binary alpha ask for the quantize step
    (Context.cpp:519-541), which this path does not run: UnsupportedFeature,
    nothing delivered."""
    img = synth.noise_rect_rgba(16, 16, seed=2)
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], fmt, N.FILTER_BOX, driver=C_DRIVER,
                                                  extra=(1, dither, binary))
    assert not ok and errors == [UNSUPPORTED_FEATURE] and stream == b"" and len(events) == 0


def test_c_interface_dithering_ignored_for_bc4(tmp_path, oracle):
    """Colour dithering applies to BC1-BC3 only (Context.cpp:521-526): BC4 with
    it set compresses as without."""
    img = synth.single_channel_rgba(32, 32, seed=3)
    a = run_process(tmp_path, [[img]], N.FORMAT_BC4, N.FILTER_BOX, driver=C_DRIVER, extra=(1, 1, 0))
    b = run_process(tmp_path, [[img]], N.FORMAT_BC4, N.FILTER_BOX, driver=C_DRIVER, extra=(1, 0, 0))
    assert a[3] and a[2] == [] and a[0] == b[0]


def test_c_interface_file_output(tmp_path):
    """nvttSetOutputOptionsFileName: the file holds exactly the stream the
    C++ front end hands its OutputHandler (header and every level)."""
    img = synth.noise_rect_rgba(72, 40, seed=61, transparent=True)
    out = tmp_path / "c.dds"
    stream, events, errors, ok, est = run_process(tmp_path, [[img]], N.FORMAT_BC3, N.FILTER_KAISER, driver=C_DRIVER,
                                                  extra=(1, 0, 0, str(out)))
    assert ok and errors == [] and stream == b"" and len(events) == 0
    ref = run_process(tmp_path, [[img]], N.FORMAT_BC3, N.FILTER_KAISER)
    data = out.read_bytes()
    # the C++ driver sets user version 7 (reserved[8], byte 64); the C interface has no setter
    assert data[:64] == ref[0][:64] and data[68:] == ref[0][68:] and data[64:68] == bytes(4)
