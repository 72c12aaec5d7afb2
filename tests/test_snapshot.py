"""Native snapshot files (hy_snapshot_write / hy_snapshot_read) without a GPU.

The schema is the reference's single-file HDF5 layout
(include/cosmo_nbody/io/output_schema.hpp:22-140, src/io/snapshot_io_write_impl.inc).
Files are read back through the library and, independently, through the HDF5
C library called directly here (attributes, shapes, dataset units and values);
the reference reader itself cannot be built in this image (no HDF5 headers),
so its checks are restated by the library's reader and tested for each
failure it raises."""
import ctypes as C
import os

import numpy as np
import pytest


def _hdf5():
    for name in (os.environ.get("HY_HDF5_LIBRARY"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"):
        if not name:
            continue
        try:
            return C.CDLL(name)
        except OSError:
            continue
    pytest.skip("no HDF5 C library in this image")


@pytest.fixture(scope="module")
def snap():
    import hyowon_amd  # noqa: F401  (builds / loads the library)
    from hyowon_amd import snapshot
    _hdf5()
    return snapshot


def _state(n, L, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, L, (3, n))
    mom = rng.normal(0.0, 300.0, (3, n))
    return pos, mom


def _meta(snap, L=64.0, n_side=16):
    return snap.SnapshotMeta(box_size=L, particles_per_dim=n_side, pm_mesh_per_dim=2 * n_side, ic_mesh_per_dim=n_side,
                             seed=7, delta_ln_a=0.1)


class RawH5:
    """Reads of a snapshot through the HDF5 C API directly (independent of the library's reader)."""

    def __init__(self, path):
        H = self.H = _hdf5()
        hid = C.c_int64
        for name, res, args in [("H5open", C.c_int, []), ("H5Fopen", hid, [C.c_char_p, C.c_uint, hid]),
                                ("H5Fclose", C.c_int, [hid]), ("H5Aopen_by_name", hid, [hid, C.c_char_p, C.c_char_p, hid, hid]),
                                ("H5Aread", C.c_int, [hid, hid, C.c_void_p]), ("H5Aclose", C.c_int, [hid]),
                                ("H5Aget_type", hid, [hid]), ("H5Tget_size", C.c_size_t, [hid]),
                                ("H5Dopen2", hid, [hid, C.c_char_p, hid]), ("H5Dread", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                                ("H5Dclose", C.c_int, [hid]), ("H5Dget_space", hid, [hid]),
                                ("H5Sget_simple_extent_dims", C.c_int, [hid, C.POINTER(C.c_ulonglong), C.c_void_p]),
                                ("H5Sclose", C.c_int, [hid]), ("H5Lexists", C.c_int, [hid, C.c_char_p, hid])]:
            f = getattr(H, name)
            f.restype = res
            f.argtypes = args
        H.H5open()
        self.dbl = C.c_int64.in_dll(H, "H5T_NATIVE_DOUBLE_g").value
        self.u32 = C.c_int64.in_dll(H, "H5T_NATIVE_UINT32_g").value
        self.u64 = C.c_int64.in_dll(H, "H5T_NATIVE_UINT64_g").value
        self.i32 = C.c_int64.in_dll(H, "H5T_NATIVE_INT32_g").value
        self.f = H.H5Fopen(path.encode(), 0, 0)
        assert self.f >= 0

    def close(self):
        self.H.H5Fclose(self.f)

    def exists(self, path):
        return self.H.H5Lexists(self.f, path.encode(), 0) > 0

    def attr(self, obj, name, ctype, count=1):
        a = self.H.H5Aopen_by_name(self.f, obj.encode(), name.encode(), 0, 0)
        assert a >= 0, (obj, name)
        buf = (ctype * count)()
        t = {C.c_double: self.dbl, C.c_uint32: self.u32, C.c_uint64: self.u64, C.c_int32: self.i32}[ctype]
        assert self.H.H5Aread(a, t, buf) >= 0
        self.H.H5Aclose(a)
        return list(buf) if count > 1 else buf[0]

    def string_attr(self, obj, name):
        a = self.H.H5Aopen_by_name(self.f, obj.encode(), name.encode(), 0, 0)
        assert a >= 0, (obj, name)
        t = self.H.H5Aget_type(a)
        size = self.H.H5Tget_size(t)
        buf = C.create_string_buffer(size + 1)
        assert self.H.H5Aread(a, t, buf) >= 0
        self.H.H5Aclose(a)
        return buf.value.decode()

    def dataset(self, path, dtype, shape):
        d = self.H.H5Dopen2(self.f, path.encode(), 0)
        assert d >= 0, path
        sp = self.H.H5Dget_space(d)
        dims = (C.c_ulonglong * 2)()
        rank = self.H.H5Sget_simple_extent_dims(sp, dims, None)
        self.H.H5Sclose(sp)
        assert tuple(dims[:rank]) == shape
        out = np.zeros(shape, dtype)
        t = self.dbl if dtype == np.float64 else self.u64
        assert self.H.H5Dread(d, t, 0, 0, 0, out.ctypes.data_as(C.c_void_p)) >= 0
        self.H.H5Dclose(d)
        return out


def test_round_trip_uniform_mass(snap, tmp_path):
    n, L, a = 4096, 64.0, 0.37
    pos, mom = _state(n, L, 1)
    ids = np.arange(n, dtype=np.uint64)[::-1].copy() + 100
    meta = _meta(snap, L)
    path = str(tmp_path / "snap_000.hdf5")
    snap.write(path, meta, a, *pos, *mom, ids=ids, uniform_mass=2.5)
    got = snap.read(path, expect=meta)
    assert got["a"] == a and got["uniform_mass"] == 2.5 and got["masses"] is None
    assert np.array_equal(got["positions"], pos)
    assert np.array_equal(got["ids"], ids)
    # velocities are stored as p / a and read back as v a (the reference's conversions)
    assert np.array_equal(got["momenta"], (mom * (1.0 / a)) * a)
    assert np.allclose(got["momenta"], mom, rtol=1e-15, atol=0)


def test_round_trip_explicit_masses_and_default_ids(snap, tmp_path):
    n, L, a = 1000, 32.0, 1.0
    pos, mom = _state(n, L, 2)
    masses = np.random.default_rng(3).uniform(0.5, 2.0, n)
    meta = _meta(snap, L)
    path = str(tmp_path / "snap_m.hdf5")
    snap.write(path, meta, a, *pos, *mom, masses=masses)
    got = snap.read(path)
    assert np.array_equal(got["masses"], masses)
    assert np.array_equal(got["ids"], np.arange(n, dtype=np.uint64))
    assert got["uniform_mass"] is None


def test_schema_through_the_hdf5_api(snap, tmp_path):
    """Header, Parameters, Config and PartType1 as the reference writes them."""
    n, L, a = 512, 64.0, 0.5
    pos, mom = _state(n, L, 4)
    meta = _meta(snap, L)
    path = str(tmp_path / "snap_raw.hdf5")
    snap.write(path, meta, a, *pos, *mom, uniform_mass=1.75)
    r = RawH5(path)
    try:
        assert r.attr("/Header", "NumPart_ThisFile", C.c_uint32, 6) == [0, n, 0, 0, 0, 0]
        assert r.attr("/Header", "NumPart_Total", C.c_uint32, 6) == [0, n, 0, 0, 0, 0]
        assert r.attr("/Header", "MassTable", C.c_double, 6) == [0, 1.75, 0, 0, 0, 0]
        assert r.attr("/Header", "Time", C.c_double) == a
        assert r.attr("/Header", "Redshift", C.c_double) == 1.0 / a - 1.0
        assert r.attr("/Header", "BoxSize", C.c_double) == L
        assert r.attr("/Header", "Omega0", C.c_double) == 0.3
        assert r.attr("/Header", "OmegaLambda", C.c_double) == 0.7
        assert r.attr("/Header", "HubbleParam", C.c_double) == 0.7
        assert r.attr("/Header", "NumFilesPerSnapshot", C.c_int32) == 1
        assert r.attr("/Header", "Flag_DoublePrecision", C.c_int32) == 1
        assert r.attr("/Header", "UnitLength_in_cm", C.c_double) == 3.085677581491367e22 * 100.0
        assert r.attr("/Header", "UnitMass_in_g", C.c_double) == 1e10 * 1.98847e30 * 1000.0
        assert r.attr("/Header", "UnitVelocity_in_cm_per_s", C.c_double) == 1e5
        assert r.attr("/Parameters", "Seed", C.c_uint64) == 7
        assert r.attr("/Parameters", "ParticlesPerDim", C.c_uint32) == 16
        assert r.attr("/Parameters", "PMMeshPerDim", C.c_uint32) == 32
        assert r.string_attr("/Parameters", "ICAmplitudeMode") == "gaussian"
        assert r.string_attr("/Config", "SoftwareName") == "HYOWON"
        assert r.string_attr("/Config", "SnapshotSchema") == "hyowon.snapshot"
        assert r.string_attr("/Config", "ParticleSpecies") == "cold_plus_baryon_collisionless_mnu0"
        assert r.string_attr("/Config", "PeriodicBoundaryConvention") == "wrapped_half_open_0_L"
        fp = r.string_attr("/Config", "PhysicsFingerprint")
        assert fp.startswith("snapshot_physics\nh=0.69999999999999996\n") and "gravity_solver=2:PM\n" in fp
        assert r.string_attr("/Config", "RunMetadataJson").startswith("{")
        coords = r.dataset("/PartType1/Coordinates", np.float64, (n, 3))
        assert np.array_equal(coords, pos.T)
        vel = r.dataset("/PartType1/Velocities", np.float64, (n, 3))
        assert np.array_equal(vel, (mom * (1.0 / a)).T)
        assert np.array_equal(r.dataset("/PartType1/ParticleIDs", np.uint64, (n,)), np.arange(n, dtype=np.uint64))
        assert not r.exists("/PartType1/Masses")
    finally:
        r.close()


def test_writer_and_reader_fail_closed(snap, tmp_path):
    from hyowon_amd._lib import HyInvalidArgument, HyRuntimeError
    n, L = 64, 16.0
    pos, mom = _state(n, L, 5)
    meta = _meta(snap, L, 4)
    path = str(tmp_path / "s.hdf5")
    snap.write(path, meta, 0.5, *pos, *mom, uniform_mass=1.0)
    with pytest.raises(HyRuntimeError):  # require-absent publication
        snap.write(path, meta, 0.5, *pos, *mom, uniform_mass=1.0)
    snap.write(path, meta, 0.25, *pos, *mom, uniform_mass=1.0, replace=True)
    assert snap.read(path)["a"] == 0.25
    bad = pos.copy()
    bad[0, 3] = L  # outside [0, L)
    with pytest.raises(HyInvalidArgument):
        snap.write(str(tmp_path / "b.hdf5"), meta, 0.5, *bad, *mom, uniform_mass=1.0)
    with pytest.raises(HyInvalidArgument):
        snap.write(str(tmp_path / "c.hdf5"), meta, 0.5, *pos, *mom)  # no mass at all
    other = snap.SnapshotMeta(box_size=2 * L, particles_per_dim=4, pm_mesh_per_dim=8)
    with pytest.raises(HyRuntimeError, match="disagrees with the admitted read context"):
        snap.read(path, expect=other)
    with pytest.raises(HyRuntimeError):
        snap.read(str(tmp_path / "missing.hdf5"))
    empty = str(tmp_path / "e.hdf5")
    snap.write(empty, meta, 1.0, *(np.zeros(0) for _ in range(6)), uniform_mass=1.0)
    got = snap.read(empty)
    assert got["positions"].shape == (3, 0)
