"""".hpcs" checkpoint files (CPU): the byte layout, files written by the
reference's own checkpoint code (src/c/checkpoint/checkpoint.c, compiled into
oracle/_ref) read back here and the other way round, and the reference's
failure statuses (tests/regression/checkpoint_test.c covers the same three
behaviours: round trip, peek, garbage rejected)."""
import ctypes as C
import struct

import numpy as np
import pytest

from nbx_testlib import Error, View

OK, INVALID_ARGUMENT, INVALID_STATE, IO_ERROR = 0, 1, 3, 7


@pytest.fixture(scope="module")
def ckpt():
    from n_body_sim_pro_amd import checkpoint
    return checkpoint


def _state(n, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(n) for k in ("px", "py", "pz", "vx", "vy", "vz",
                                                 "ax", "ay", "az")} | {"m": rng.uniform(0.1, 1, n)}


def _hdr(ckpt, n):
    return ckpt.header(n, simulation_time=1.25, timestep=5e-4, integrator=2, theta=0.6,
                       barnes_hut=True, seed=0xDEADBEEF12345678, preset=7)


def test_layout(ckpt, tmp_path):
    n = 37
    s = _state(n, 1)
    path = tmp_path / "a.hpcs"
    ckpt.write(path, s, _hdr(ckpt, n))
    raw = path.read_bytes()
    assert len(raw) == 80 + 10 * 8 * n
    magic, version, hmagic, hversion, count = struct.unpack_from("<IIIIQ", raw, 0)
    assert (magic, version, hmagic, hversion, count) == (0x53504348, 1, 0x53504348, 1, n)
    assert raw[0:4] == b"HCPS"        # the constant "HPCS" (0x53504348) stored little-endian
    t, dt, integ = struct.unpack_from("<ddi", raw, 24)
    theta, bh = struct.unpack_from("<di", raw, 48)
    seed, preset = struct.unpack_from("<Qi", raw, 64)
    assert (t, dt, integ, theta, bh, seed, preset) == (1.25, 5e-4, 2, 0.6, 1, 0xDEADBEEF12345678, 7)
    assert raw[44:48] == raw[60:64] == raw[76:80] == b"\0\0\0\0"      # padding
    body = np.frombuffer(raw[80:], dtype="<f8").reshape(10, n)
    for k, name in enumerate(ckpt.ARRAY_NAMES):
        assert np.array_equal(body[k], s[name])


def test_round_trip_and_peek(ckpt, tmp_path):
    n = 1000
    s = _state(n, 2)
    path = tmp_path / "b.hpcs"
    ckpt.write(path, s, _hdr(ckpt, n))
    p = ckpt.peek(path)
    assert (p.particle_count, p.simulation_time, p.random_seed, p.preset) == (
        n, 1.25, 0xDEADBEEF12345678, 7)
    h, a = ckpt.read(path, capacity=n + 5)
    assert h.particle_count == n and h.integrator == 2 and h.theta == 0.6
    for k in ckpt.ARRAY_NAMES:
        assert np.array_equal(a[k], s[k])
    # zero particles is a valid file
    empty = tmp_path / "empty.hpcs"
    ckpt.write(empty, {k: np.zeros(0) for k in ckpt.ARRAY_NAMES}, _hdr(ckpt, 0))
    assert empty.stat().st_size == 80
    assert ckpt.read(empty, capacity=0)[0].particle_count == 0


def _status(excinfo):
    return excinfo.value.status


def test_failures(ckpt, tmp_path):
    from n_body_sim_pro_amd import NBXError
    n = 16
    s = _state(n, 3)
    good = tmp_path / "good.hpcs"
    ckpt.write(good, s, _hdr(ckpt, n))
    garbage = tmp_path / "garbage.hpcs"
    garbage.write_bytes(b"this is not a checkpoint")
    with pytest.raises(NBXError) as e:
        ckpt.peek(garbage)
    assert _status(e) == INVALID_STATE
    wrong_version = tmp_path / "v2.hpcs"
    raw = bytearray(good.read_bytes())
    raw[4:8] = struct.pack("<I", 2)
    wrong_version.write_bytes(bytes(raw))
    with pytest.raises(NBXError) as e:
        ckpt.read(wrong_version)
    assert _status(e) == INVALID_STATE
    truncated = tmp_path / "short.hpcs"
    truncated.write_bytes(good.read_bytes()[:-8])
    with pytest.raises(NBXError) as e:
        ckpt.read(truncated)
    assert _status(e) == IO_ERROR
    tiny = tmp_path / "tiny.hpcs"
    tiny.write_bytes(good.read_bytes()[:6])
    with pytest.raises(NBXError) as e:
        ckpt.peek(tiny)
    assert _status(e) == IO_ERROR
    with pytest.raises(NBXError) as e:
        ckpt.read(good, capacity=n - 1)
    assert _status(e) == INVALID_ARGUMENT
    with pytest.raises(NBXError) as e:
        ckpt.write(tmp_path / "x.hpcs", s, _hdr(ckpt, n + 1))   # count differs from the view
    assert _status(e) == INVALID_ARGUMENT
    with pytest.raises(NBXError) as e:
        ckpt.peek(tmp_path / "missing.hpcs")
    assert _status(e) == IO_ERROR


@pytest.fixture(scope="module")
def refckpt(ref):
    hp = C.c_void_p
    ref.n_body_sim_pro_checkpoint_write.argtypes = [C.c_char_p, C.POINTER(View), hp,
                                                    C.POINTER(Error)]
    ref.n_body_sim_pro_checkpoint_read.argtypes = [C.c_char_p, hp, C.c_void_p, C.POINTER(Error)]
    ref.n_body_sim_pro_checkpoint_peek.argtypes = [C.c_char_p, hp, C.POINTER(Error)]
    for f in ("write", "read", "peek"):
        getattr(ref, f"n_body_sim_pro_checkpoint_{f}").restype = C.c_int
    return ref


def test_reference_writes_the_same_bytes(ckpt, refckpt, tmp_path):
    n = 513
    s = _state(n, 4)
    ours, theirs = tmp_path / "ours.hpcs", tmp_path / "ref.hpcs"
    hdr = _hdr(ckpt, n)
    ckpt.write(ours, s, hdr)
    from n_body_sim_pro_amd import make_view
    a = {k: np.ascontiguousarray(s[k]) for k in ckpt.ARRAY_NAMES}
    v = make_view(n, *(a[k] for k in ckpt.ARRAY_NAMES))
    err = Error()
    assert refckpt.n_body_sim_pro_checkpoint_write(str(theirs).encode(), C.byref(v),
                                                   C.cast(C.byref(hdr), C.c_void_p),
                                                   C.byref(err)) == OK
    assert ours.read_bytes() == theirs.read_bytes()


def test_cross_reads(ckpt, refckpt, tmp_path):
    n = 777
    s = _state(n, 5)
    theirs, ours = tmp_path / "ref.hpcs", tmp_path / "ours.hpcs"
    from n_body_sim_pro_amd import make_view
    a = {k: np.ascontiguousarray(s[k]) for k in ckpt.ARRAY_NAMES}
    v = make_view(n, *(a[k] for k in ckpt.ARRAY_NAMES))
    hdr = _hdr(ckpt, n)
    err = Error()
    assert refckpt.n_body_sim_pro_checkpoint_write(str(theirs).encode(), C.byref(v),
                                                   C.cast(C.byref(hdr), C.c_void_p),
                                                   C.byref(err)) == OK
    # reference file -> this reader
    h, got = ckpt.read(theirs)
    assert (h.particle_count, h.timestep, h.preset) == (n, 5e-4, 7)
    for k in ckpt.ARRAY_NAMES:
        assert np.array_equal(got[k], s[k])
    # this writer -> reference reader (into its own particle system)
    ckpt.write(ours, s, hdr)
    ps = refckpt.n_body_sim_pro_particle_system_create(n)
    try:
        rh = ckpt.CheckpointHeader()
        assert refckpt.n_body_sim_pro_checkpoint_read(str(ours).encode(),
                                                      C.cast(C.byref(rh), C.c_void_p), ps,
                                                      C.byref(err)) == OK
        assert (rh.particle_count, rh.random_seed, rh.theta) == (n, 0xDEADBEEF12345678, 0.6)
        rv = View()
        assert refckpt.n_body_sim_pro_particle_system_view(ps, C.byref(rv), C.byref(err)) == OK
        fields = ("positions_x", "positions_y", "positions_z", "velocities_x", "velocities_y",
                  "velocities_z", "accelerations_x", "accelerations_y", "accelerations_z",
                  "masses")
        for name, f in zip(ckpt.ARRAY_NAMES, fields):
            arr = np.ctypeslib.as_array(getattr(rv, f), shape=(n,))
            assert np.array_equal(arr, s[name])
    finally:
        refckpt.n_body_sim_pro_particle_system_destroy(ps)


@pytest.mark.gpu
def test_engine_save_load_round_trip(ckpt, nbx, tmp_path):
    """Device state -> file -> another engine: identical arrays, and both
    engines then step identically (Barnes-Hut, which rebuilds its tree from
    the loaded state)."""
    n = 50000
    p = nbx.two_galaxy(n, seed=17)
    path = tmp_path / "state.hpcs"
    with nbx.Engine(n) as a:
        a.upload(p["pos"], p["m"], vel=p["vel"])
        a.compute_force("hip_barnes_hut")
        a.advance("hip_barnes_hut", dt=1e-3, steps=2)
        hdr = ckpt.header(0, simulation_time=0.002, timestep=1e-3, integrator=1, theta=0.7,
                          barnes_hut=True, seed=17, preset=7)
        ckpt.save_engine(a, path, hdr)
        before = a.download()
        a.advance("hip_barnes_hut", dt=1e-3, steps=3)
        after_a = a.download()
    h, arrays = ckpt.read(path)
    assert h.particle_count == n and h.simulation_time == 0.002 and h.barnes_hut_enabled == 1
    assert np.array_equal(np.stack([arrays["px"], arrays["py"], arrays["pz"]], 1), before["pos"])
    assert np.array_equal(np.stack([arrays["ax"], arrays["ay"], arrays["az"]], 1), before["acc"])
    from n_body_sim_pro_amd import NBXError
    with nbx.Engine(n - 1) as small, pytest.raises(NBXError) as e:
        ckpt.load_engine(small, path)
    assert e.value.status == INVALID_ARGUMENT
    with nbx.Engine(n + 100) as b:
        got = ckpt.load_engine(b, path)
        assert got.particle_count == n and got.random_seed == 17
        loaded = b.download()
        for k in ("pos", "vel", "acc", "m"):
            assert np.array_equal(loaded[k], before[k]), k
        b.advance("hip_barnes_hut", dt=1e-3, steps=3)
        after_b = b.download()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(after_a[k], after_b[k]), k
