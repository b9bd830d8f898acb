"""Ping files (include/rtsim_sink.h): the FileSink drop-in writes a stream of
ping records; the byte layout is checked independently with numpy, the
readers round-trip, corruption is detected, and (GPU) a computed ping is
written and read back bit for bit."""
import numpy as np
import pytest

import rtac_simulation_amd as R
from rtac_simulation_amd import configs


def _ping(H, W, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))).astype(np.complex64)


def test_round_trip_and_layout(tmp_path):
    path = tmp_path / "a" / "b" / "pings.rtpg"
    sink = R.FileSink.Create(path)
    p2 = _ping(5, 3, 0)
    b2 = np.array([-0.4, 0.0, 0.4], np.float32)
    p1 = _ping(7, 1, 1)[:, 0]
    sink.write_host(2, p2, 0.1, 30.0, b2)
    sink.write_host(1, p1, 0.0, 50.0)
    assert sink.count() == 2
    sink.close()
    assert R.ping_file_count(path) == 2
    r = R.read_ping(path, 0)
    assert (r["kind"], r["height"], r["width"], r["index"]) == (2, 5, 3, 0)
    assert np.array_equal(r["data"], p2) and np.array_equal(r["bearings"], b2)
    r = R.read_ping(path, 1)
    assert (r["kind"], r["height"], r["width"], r["index"]) == (1, 7, 1, 1)
    assert np.array_equal(r["data"], p1) and r["bearings"].size == 0
    # independent parse of the documented layout
    raw = path.read_bytes()
    assert raw[:4] == b"RTPG"
    v, kind, scalar, H, W = np.frombuffer(raw[4:24], "<u4")
    lo, hi = np.frombuffer(raw[24:32], "<f4")
    idx, chk = np.frombuffer(raw[32:48], "<u8")
    assert (v, kind, scalar, H, W, idx) == (1, 2, 1, 5, 3, 0)
    assert (lo, hi) == (np.float32(0.1), np.float32(30.0))
    body = raw[48:48 + 4 * 3 + 8 * 15]
    h = np.uint64(0xcbf29ce484222325)
    for byte in body:     # FNV-1a 64
        h = np.uint64((int(h) ^ byte) * 0x100000001b3 % (1 << 64))
    assert int(h) == int(chk)
    assert np.array_equal(np.frombuffer(body[12:], "<c8").reshape(5, 3), p2)
    second = 48 + len(body)
    assert raw[second:second + 4] == b"RTPG"
    assert len(raw) == second + 48 + 8 * 7


def test_existing_file_and_corruption(tmp_path):
    path = tmp_path / "pings.rtpg"
    s = R.FileSink.Create(path)
    s.write_host(2, _ping(4, 4, 2), 0.1, 30.0, np.linspace(-1, 1, 4, dtype=np.float32))
    s.close()
    with pytest.raises(R.RtsimError) as e:
        R.FileSink.Create(path)
    assert e.value.code == R.ERR_IO
    raw = bytearray(path.read_bytes())
    raw[60] ^= 0x01
    path.write_bytes(bytes(raw))
    with pytest.raises(R.RtsimError) as e:
        R.read_ping(path, 0)
    assert e.value.code == R.ERR_IO and "checksum" in str(e.value)
    s = R.FileSink.Create(path, overwrite=True)
    s.close()
    assert R.ping_file_count(path) == 0


def test_truncated_and_oversized_records_are_io_errors(tmp_path):
    """A header is checked against the file before anything is sized from it:
    a truncated last record, a header announcing 2^32 x 2^32 cells, an
    unknown kind and a 1D record wider than 1 are I/O errors, never a crash."""
    path = tmp_path / "pings.rtpg"
    s = R.FileSink.Create(path)
    s.write_host(2, _ping(6, 4, 3), 0.1, 30.0, np.linspace(-1, 1, 4, dtype=np.float32))
    s.write_host(2, _ping(6, 4, 4), 0.1, 30.0, np.linspace(-1, 1, 4, dtype=np.float32))
    s.close()
    raw = path.read_bytes()
    rec = len(raw) // 2
    (tmp_path / "cut.rtpg").write_bytes(raw[:-5])
    with pytest.raises(R.RtsimError) as e:
        R.ping_file_count(tmp_path / "cut.rtpg")
    assert e.value.code == R.ERR_IO and "truncated" in str(e.value)
    assert np.array_equal(R.read_ping(tmp_path / "cut.rtpg", 0)["data"], _ping(6, 4, 3))

    def patched(offset, value):
        b = bytearray(raw[:rec])
        b[offset:offset + 4] = np.uint32(value).tobytes()
        q = tmp_path / f"bad_{offset}_{value}.rtpg"
        q.write_bytes(bytes(b))
        return q

    for off, val, msg in [(16, 0xFFFFFFFF, "truncated"), (20, 0xFFFFFFFF, "truncated"),
                          (8, 7, "kind"), (8, 1, "kind")]:
        q = patched(off, val)
        for fn in (lambda: R.ping_file_count(q), lambda: R.read_ping(q, 0)):
            with pytest.raises(R.RtsimError) as e:
                fn()
            assert e.value.code == R.ERR_IO and msg in str(e.value), (off, val, str(e.value))
    big = bytearray(raw[:rec])
    big[16:20] = np.uint32(0xFFFFFFFF).tobytes()
    big[20:24] = np.uint32(0xFFFFFFFF).tobytes()
    (tmp_path / "huge.rtpg").write_bytes(bytes(big))
    with pytest.raises(R.RtsimError) as e:
        R.read_ping(tmp_path / "huge.rtpg", 0)
    assert e.value.code == R.ERR_IO


@pytest.mark.gpu
def test_sink_writes_computed_pings(gpu, tmp_path):
    R_ = gpu
    path = tmp_path / "fls.rtpg"
    sink = R_.FileSink.Create(path)
    cfg = configs._fls(48, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "sink")
    s = R_.make_sensor(cfg)
    s.set_sample_count(20_000)
    pings = []
    for k, form in enumerate((0, 2)):
        s.write_samples(R_.generate_hits(configs.fls_scene(), 70 + k, 20_000).to_device())
        s.compute_output(form)
        sink.set_output(s)
        pings.append(s.output().copy())
    c1, sc1, _, _ = configs.workload("single_beam_38k")
    s1 = R_.make_sensor(c1)
    s1.set_sample_count(5000)
    s1.write_samples(R_.generate_hits(sc1, 3, 5000).to_device())
    s1.compute_output(0)
    sink.set_output(s1)
    sink.close()
    assert R_.ping_file_count(path) == 3
    for k in range(2):
        r = R_.read_ping(path, k)
        assert r["kind"] == 2 and r["height"] == cfg.H and r["width"] == cfg.W
        assert np.array_equal(r["data"].view(np.uint32), pings[k].view(np.uint32))
        assert np.array_equal(r["bearings"], np.asarray(cfg.bearings, np.float32))
    r = R_.read_ping(path, 2)
    assert r["kind"] == 1 and np.array_equal(r["data"], s1.output())
