"""cmpmc_amd: MI355X (gfx950) implementation of CMPMC's track-transport path.

Python front end over the C ABI (include/cmpmc_amd.h, library cmpmc_amd/_build/
libcmpmc_amd.so).  The reference's C++ API has the same shape:

    CrystalLattice(cfg)            -> Lattice(crystal)            (host table build)
    DetectorModel{}                -> default_detector()
    SimulationEngine(lat, det)     -> Engine(lattice, detector)   (tables uploaded to HBM)
    engine.transport(stack, ...)   -> Engine.transport(tracks, rng_state)
    run_events(engine, n, gen, o)  -> Engine.run_events(batch, ...) (host-generated stacks)
                                      Engine.run_workload(wl, n, ...) (device generator)

Every compute call goes to the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _capi as capi
from ._capi import (CHARGE_DTYPE, FORM_BATCHED, FORM_PLAIN, HIT_DTYPE, MODE_BENCH_DOS, NIEL_EMPIRICAL,
                    NIEL_IMPACT_TUNL, NIEL_LEWIN_SMITH, NIEL_LINDHARD, NIEL_SARKIS, PHONON_DTYPE, WL_DEPOSIT, WL_PAIR,
                    WL_PHONON_LOGN, WL_PHONON_POINT, WL_PHONON_UNIF, CmpmcError, check)

__all__ = ["Lattice", "Engine", "DiffTape", "apply_diff_params", "Rng", "germanium", "default_detector", "set_bias", "demo_polynomial_surfaces",
           "workload", "make_tracks", "CmpmcError", "FORM_PLAIN", "FORM_BATCHED", "PHONON_DTYPE", "CHARGE_DTYPE",
           "HIT_DTYPE", "WL_PHONON_POINT", "WL_PHONON_LOGN", "WL_PAIR", "WL_PHONON_UNIF", "MODE_BENCH_DOS", "device_info",
           "constant_split_surfaces", "set_pads", "clear_pads", "qet_mask", "grid_mask", "find_channel", "medium",
           "medium_from_bulk_shear", "medium_from_lame", "set_amm", "amm_transmission", "amm_psv", "WL_DEPOSIT",
           "partition_opts", "ionization_yield", "lindhard_yield", "sample_fano_binomial", "partition_deposit",
           "NIEL_LINDHARD", "NIEL_LEWIN_SMITH", "NIEL_SARKIS", "NIEL_IMPACT_TUNL", "NIEL_EMPIRICAL", "Mesh",
           "set_mesh"]

MASK64 = (1 << 64) - 1


# ------------------------------------------------------------------ host RNG (core/rng.hpp)
def _mix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


class Rng:
    """xoshiro256** with splitmix64 seeding and Rng::for_stream (rng.hpp:17-97), for
    host-side generators that hand post-generator stream states to the device."""

    def __init__(self, seed: int = 0x853C49E6748FEA9B):
        z = seed & MASK64
        s = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & MASK64
            x = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
            s.append(x ^ (x >> 31))
        self.s = s

    @classmethod
    def for_stream(cls, seed: int, stream: int) -> "Rng":
        return cls(_mix64(seed & MASK64) ^ _mix64((stream + 0x632BE59BD9B4E019) & MASK64))

    def next_u64(self) -> int:
        s = self.s
        r = ((((s[1] * 5) & MASK64) << 7 | ((s[1] * 5) & MASK64) >> 57) & MASK64) * 9 & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & MASK64
        return r

    def uniform(self) -> float:
        return float((self.next_u64() >> 11) | 1) * 2.0 ** -53

    def isotropic_dir(self) -> tuple:
        ct = 2.0 * self.uniform() - 1.0
        st = max(0.0, 1.0 - ct * ct) ** 0.5
        phi = 2.0 * 3.141592653589793 * self.uniform()
        return (st * np.cos(phi), st * np.sin(phi), ct)

    def state(self) -> np.ndarray:
        return np.array(self.s, dtype=np.uint64)


# ------------------------------------------------------------------ configuration
def device_info() -> str:
    buf = C.create_string_buffer(256)
    check(capi.load().cmpmc_device_info(buf, 256))
    return buf.value.decode()


def germanium() -> capi.Crystal:
    c = capi.Crystal()
    check(capi.load().cmpmc_crystal_germanium(C.byref(c)))
    return c


def default_detector() -> capi.Detector:
    d = capi.Detector()
    check(capi.load().cmpmc_detector_default(C.byref(d)))
    return d


def aluminum_film() -> capi.Film:
    """SuperconductorFilmConfig of the Al sensor film (crystal_config.hpp:240-257)."""
    f = capi.Film()
    check(capi.load().cmpmc_film_aluminum(C.byref(f)))
    return f


def _film_call(fn, energy: float, rng: "Rng", *args) -> tuple[float, np.ndarray]:
    st = (capi.u64 * 4)(*rng.s)
    refl = np.zeros(256)
    n = capi.i32()
    edep = C.c_double()
    check(fn(*args[:1], energy, *args[1:], st, refl.ctypes.data, C.byref(n), C.byref(edep)))
    rng.s = [int(x) for x in st]
    return edep.value, refl[:n.value].copy()


def film_absorb(film: capi.Film, energy: float, rng: "Rng", photon: bool = False) -> tuple[float, np.ndarray]:
    """KaplanFilm::absorb_phonon (photon=False) / absorb_quantum with a photon seed on the
    host (kaplan_film.hpp:106-206): (absorbed energy, re-emitted energies); advances rng."""
    return _film_call(capi.load().cmpmc_film_absorb_quantum, energy, rng, C.byref(film), int(photon))


def film_loss_probability(film: capi.Film, energy: float, thickness_frac: float = 1.0) -> tuple[float, float]:
    """KaplanFilm::loss_probability and kaplan_pair_breaking_rate (kaplan_film.hpp:69-99)."""
    p, g = C.c_double(), C.c_double()
    check(capi.load().cmpmc_film_loss_probability(C.byref(film), energy, thickness_frac, C.byref(p), C.byref(g)))
    return p.value, g.value


class KaplanSurrogate:
    """KaplanSurrogateLUT(film, seed) (kaplan_film.hpp:325-383): pre-sampled cascade
    outcomes on [2 Delta, 40 Delta]; absorb() is absorb_phonon_fast."""

    def __init__(self, film: capi.Film | None = None, seed: int = 1337):
        self.film = film if film is not None else aluminum_film()
        self.ptr = C.c_void_p()
        check(capi.load().cmpmc_surrogate_create(C.byref(self.film), seed, C.byref(self.ptr)))

    def absorb(self, energy: float, rng: "Rng") -> tuple[float, np.ndarray]:
        return _film_call(capi.load().cmpmc_surrogate_absorb, energy, rng, self.ptr)

    def tables(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        a, o, r = C.POINTER(C.c_float)(), C.POINTER(capi.u32)(), C.POINTER(C.c_float)()
        n = capi.u64()
        check(capi.load().cmpmc_surrogate_tables(self.ptr, C.byref(a), C.byref(o), C.byref(r), C.byref(n)))
        k = CMPMC_SURR_NE * CMPMC_SURR_NS
        return (np.ctypeslib.as_array(a, shape=(k,)).copy(), np.ctypeslib.as_array(o, shape=(k + 1,)).copy(),
                np.ctypeslib.as_array(r, shape=(max(n.value, 1),))[:n.value].copy())

    def close(self) -> None:
        if getattr(self, "ptr", None):
            capi.load().cmpmc_surrogate_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CMPMC_SURR_NE, CMPMC_SURR_NS = 128, 512


def apply_diff_params(crystal: capi.Crystal, det: capi.Detector, params) -> None:
    """Writes (B_iso, A_anh, f_diffuse, p_film) into a crystal and detector the way the
    differentiable form applies them: every surface diffuse with probability f and
    specular otherwise, every film absorbing with probability p."""
    B, A, f, p = (float(v) for v in params)
    crystal.B_iso, crystal.A_anh = B, A
    for s in det.surf:
        s.n_spec, s.n_diff, s.n_anh = 1, 1, 0
        s.spec[0], s.diff[0], s.spec_prob_const = 1.0 - f, f, 1.0 - f
        if s.has_film:
            s.film.film_absorption = p


def constant_split_surfaces(det: capi.Detector, f_diffuse: float) -> None:
    """Every surface reflects diffusely with probability f_diffuse and specularly
    otherwise, at every energy (SurfaceProperty with constant spec/diff coefficients and
    no down-conversion, surface_property.hpp:23-60)."""
    for s in det.surf:
        s.n_spec, s.n_diff, s.n_anh = 1, 1, 0
        s.spec[0], s.diff[0], s.spec_prob_const = 1.0 - f_diffuse, f_diffuse, 1.0 - f_diffuse


# ------------------------------------------------------------------ electrode masks and AMM film entry
_PAD_ARRAYS: list = []     # pad arrays a surface points at (the model copies them at creation)


def _pad_array(pads) -> C.Array:
    if isinstance(pads, C.Array):
        return pads
    arr = (capi.Pad * len(pads))()
    for i, row in enumerate(pads):
        if isinstance(row, capi.Pad):
            arr[i] = row
            continue
        arr[i].shape, arr[i].channel_id = int(row[0]), int(row[1])
        for j, n in enumerate(("cx", "cy", "r_min", "r_max", "phi_min", "phi_max", "fill_fraction")):
            setattr(arr[i], n, float(row[2 + j]))
    return arr


def set_pads(surface: capi.Surface, pads) -> None:
    """Electrode mask of a surface (SurfaceProperty::electrode_mask,
    geometry/electrode_pattern.hpp:66-92): `pads` as Pad structures or (shape, channel_id,
    cx, cy, r_min, r_max, phi_min, phi_max, fill_fraction) rows. A phonon or carrier
    meeting the surface off every pad reflects as from bare crystal; hits carry the pad's
    channel id."""
    arr = _pad_array(pads)
    _PAD_ARRAYS.append(arr)
    surface.n_pads = len(arr)
    surface.pads = C.cast(arr, C.POINTER(capi.Pad))


def clear_pads(surface: capi.Surface) -> None:
    surface.n_pads = 0
    surface.pads = None


def qet_mask(radius_m: float, fill_fraction: float = 1.0) -> C.Array:
    """ElectrodePatternMask::supercdms_qet_mask (electrode_pattern.hpp:99-147)."""
    out = (capi.Pad * 5)()
    check(capi.load().cmpmc_qet_mask(radius_m, fill_fraction, out))
    return out


def grid_mask(nx: int, ny: int, pitch_x: float, pitch_y: float, pad_half_w: float, pad_half_h: float,
              fill_fraction: float = 1.0) -> C.Array:
    """ElectrodePatternMask::rectangular_grid_mask (electrode_pattern.hpp:150-176)."""
    out = (capi.Pad * (nx * ny))()
    check(capi.load().cmpmc_grid_mask(nx, ny, pitch_x, pitch_y, pad_half_w, pad_half_h, fill_fraction, out))
    return out


def find_channel(pads, x: float, y: float, rng: "Rng | None" = None) -> int:
    """ElectrodePatternMask::find_channel (electrode_pattern.hpp:82-92) on the host; a pad
    with fill_fraction < 1 draws one uniform from `rng` (advanced in place)."""
    arr = _pad_array(pads)
    st = (C.c_uint64 * 4)(*(rng.s if rng is not None else (1, 2, 3, 4)))
    ch = capi.load().cmpmc_find_channel(arr, len(arr), x, y, st)
    if rng is not None:
        rng.s = [int(v) for v in st]
    return int(ch)


def medium(rho: float, vL: float, vT: float) -> capi.Medium:
    """IsotropicMedium (interface_transmission.hpp:35-48)."""
    return capi.Medium(rho, vL, vT)


def medium_from_bulk_shear(rho: float, K: float, G: float) -> capi.Medium:
    """IsotropicMedium::from_bulk_shear (interface_transmission.hpp:41-43)."""
    return capi.Medium(rho, math.sqrt((K + 4.0 * G / 3.0) / rho), math.sqrt(G / rho))


def medium_from_lame(rho: float, lam: float, mu: float) -> capi.Medium:
    """IsotropicMedium::from_lame (interface_transmission.hpp:44-47)."""
    return capi.Medium(rho, math.sqrt((lam + 2.0 * mu) / rho), math.sqrt(mu / rho))


def set_amm(surface: capi.Surface, substrate, film) -> None:
    """AMM film entry (SurfaceProperty::amm_interface = IsotropicAMMInterface{substrate,
    film}): a phonon reaching an active part of the surface enters the film with the AMM
    transmission at its incidence angle instead of film.film_absorption."""
    surface.has_amm = 1
    surface.amm_substrate = substrate if isinstance(substrate, capi.Medium) else capi.Medium(*substrate)
    surface.amm_film = film if isinstance(film, capi.Medium) else capi.Medium(*film)


def amm_transmission(substrate, film, longitudinal: bool, cos_theta: float) -> float:
    """IsotropicAMMInterface::transmission (interface_transmission.hpp:123-129)."""
    s = substrate if isinstance(substrate, capi.Medium) else capi.Medium(*substrate)
    f = film if isinstance(film, capi.Medium) else capi.Medium(*film)
    return float(capi.load().cmpmc_amm_transmission(C.byref(s), C.byref(f), int(longitudinal), cos_theta))


def amm_psv(m1, m2, incident_P: bool, theta1: float) -> np.ndarray:
    """P-SV energy-flux coefficients (R_P, R_S, T_P, T_S) (interface_transmission.hpp:70-117)."""
    a = m1 if isinstance(m1, capi.Medium) else capi.Medium(*m1)
    b = m2 if isinstance(m2, capi.Medium) else capi.Medium(*m2)
    out = (C.c_double * 4)()
    check(capi.load().cmpmc_amm_psv(C.byref(a), C.byref(b), int(incident_P), theta1, out))
    return np.array(out[:])


# ------------------------------------------------------------------ FEM mesh fields
def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Mesh:
    """An FEM mesh field (geometry/mesh_field.hpp TriLinearMesh3D, dim 3, or
    BiLinearMesh2D, dim 2) held by the library (cmpmc_mesh). Give it to a detector with
    set_mesh: the charge step then reads its field instead of the uniform one."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        dim, ax, nv, nc = C.c_int(), C.c_int(), C.c_uint64(), C.c_uint64()
        check(capi.load().cmpmc_mesh_info(self._h, C.byref(dim), C.byref(ax), C.byref(nv), C.byref(nc)))
        self.dim, self.axisymmetric_rz = dim.value, bool(ax.value)
        self.num_vertices, self.num_cells = nv.value, nc.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            capi.load().cmpmc_mesh_destroy(h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @staticmethod
    def _made(fn, *args) -> "Mesh":
        out = C.c_void_p()
        check(fn(*args, C.byref(out)))
        return Mesh(out.value)

    @classmethod
    def tet(cls, xyz, potentials, tets, v_min_clamp: float = -1e30, v_max_clamp: float = 1e30) -> "Mesh":
        """TriLinearMesh3D(vertices, potentials, tetrahedra, V_min_clamp, V_max_clamp)."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(potentials, np.float64)
        t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        return cls._made(capi.load().cmpmc_mesh3d_create, _dp(x), len(x), _dp(v), _ip(t), len(t), v_min_clamp,
                         v_max_clamp)

    @classmethod
    def tri(cls, uv, potentials, tris, axisymmetric_rz: bool = True) -> "Mesh":
        """BiLinearMesh2D(vertices (u, v), potentials, triangles, axisymmetric_rz)."""
        x = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
        v = np.ascontiguousarray(potentials, np.float64)
        t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        return cls._made(capi.load().cmpmc_mesh2d_create, _dp(x), len(x), _dp(v), _ip(t), len(t),
                         int(axisymmetric_rz))

    @staticmethod
    def grid3d_vertices(half_x, half_y, half_z, nx, ny, nz) -> np.ndarray:
        n = (max(1, nx) + 1) * (max(1, ny) + 1) * (max(1, nz) + 1)
        out = np.zeros((n, 3))
        check(capi.load().cmpmc_mesh3d_grid_vertices(half_x, half_y, half_z, nx, ny, nz, _dp(out)))
        return out

    @classmethod
    def grid3d(cls, half_x, half_y, half_z, nx, ny, nz, potential) -> "Mesh":
        """TriLinearMesh3D::build_cartesian_grid: `potential` is a callable V(x, y, z)
        (vectorised over the vertex arrays) or the vertex potentials themselves."""
        x = cls.grid3d_vertices(half_x, half_y, half_z, nx, ny, nz)
        v = potential(x[:, 0], x[:, 1], x[:, 2]) if callable(potential) else potential
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (len(x),)))
        return cls._made(capi.load().cmpmc_mesh3d_grid, half_x, half_y, half_z, nx, ny, nz, _dp(v))

    @staticmethod
    def grid2d_vertices(u_min, u_max, v_min, v_max, nu, nv) -> np.ndarray:
        n = (max(1, nu) + 1) * (max(1, nv) + 1)
        out = np.zeros((n, 2))
        check(capi.load().cmpmc_mesh2d_grid_vertices(u_min, u_max, v_min, v_max, nu, nv, _dp(out)))
        return out

    @classmethod
    def grid2d(cls, u_min, u_max, v_min, v_max, nu, nv, potential, axisymmetric_rz: bool = True) -> "Mesh":
        """BiLinearMesh2D::build_grid: `potential` a callable V(u, v) or the vertex potentials."""
        x = cls.grid2d_vertices(u_min, u_max, v_min, v_max, nu, nv)
        v = potential(x[:, 0], x[:, 1]) if callable(potential) else potential
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (len(x),)))
        return cls._made(capi.load().cmpmc_mesh2d_grid, u_min, u_max, v_min, v_max, nu, nv, _dp(v),
                         int(axisymmetric_rz))

    @classmethod
    def load3d(cls, coords_file, tets_file, coord_unit_m: float = 1.0, potential_unit_V: float = 1.0) -> "Mesh":
        """TriLinearMesh3D::load_from_files."""
        return cls._made(capi.load().cmpmc_mesh3d_load, str(coords_file).encode(), str(tets_file).encode(),
                         coord_unit_m, potential_unit_V)

    def locate(self, p, hint: int = 0):
        """find_tetrahedron / find_triangle at mesh coordinates p: (inside, bary, hint)."""
        q = np.ascontiguousarray(p, np.float64)
        b = np.zeros(self.dim + 1)
        h = C.c_int32(hint)
        r = capi.load().cmpmc_mesh_locate(self._h, _dp(q), C.byref(h), _dp(b))
        if r < 0:
            check(r)
        return bool(r), b, h.value

    def field_at(self, p, hint: int = 0):
        """field_at: (E [V/m], hint) at a 3D position."""
        q = np.ascontiguousarray(p, np.float64)
        E = np.zeros(3)
        h = C.c_int32(hint)
        check(capi.load().cmpmc_mesh_field(self._h, _dp(q), C.byref(h), _dp(E)))
        return E, h.value

    def potential_at(self, p, hint: int = 0):
        """potential_at: (V, hint) at a 3D position."""
        q = np.ascontiguousarray(p, np.float64)
        h = C.c_int32(hint)
        v = capi.load().cmpmc_mesh_potential(self._h, _dp(q), C.byref(h))
        return float(v), h.value

    def query(self, pts, hint0: int = 0) -> dict:
        """Per point, in order, with the walk hint carried from point to point: the walk at
        the point's mesh coordinates (found, bary, the hint it leaves), field_at and
        potential_at from the same start."""
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = len(pts)
        o = {"found": np.zeros(n, np.int32), "bary": np.zeros((n, self.dim + 1)), "hints": np.zeros(n, np.int32),
             "E": np.zeros((n, 3)), "pot": np.zeros(n)}
        h = int(hint0)
        for i, p in enumerate(pts):
            if self.dim == 2:
                # the C library's hypot, as the field query's projection uses
                q = (float(np.hypot(p[0], p[1])), p[2]) if self.axisymmetric_rz else (p[0], p[1])
            else:
                q = p
            inside, b, hf = self.locate(q, h)
            o["found"][i], o["bary"][i], o["hints"][i] = int(inside), b, hf
            o["E"][i] = self.field_at(p, h)[0]
            o["pot"][i] = self.potential_at(p, h)[0]
            h = hf
        return o


def set_mesh(det: capi.Detector, mesh: "Mesh | None") -> None:
    """DetectorModel::mesh_field_3d / mesh_field_2d (None: the uniform efield). A model
    copies the mesh at creation; keep `mesh` alive until then."""
    det.mesh = None if mesh is None else mesh.handle.value
    if mesh is not None:
        _MESHES.append(mesh)


_MESHES: list = []     # meshes a detector points at (kept alive with the process)


def set_bias(det: capi.Detector, volts: float, thickness_m: float = 2.54e-2) -> None:
    """UniformElectricField::from_bias(V, d) = (0, 0, -V/d) (charge_transport.hpp:22-24)."""
    det.efield[0], det.efield[1], det.efield[2] = 0.0, 0.0, -volts / thickness_m


def demo_polynomial_surfaces(det: capi.Detector, max_bounces: int = 100) -> None:
    """The reference benchmark's surfaces (cmpmc_bench.cpp:250-258; the demonstration
    polynomials of surface_property.hpp:254-272)."""
    spec = [0.928, -2.03e-4, -3.21e-6, 3.1e-9, 2.9e-13]
    diff = [5.88e-2, 7.83e-4, -2.47e-6, 1.71e-8, -2.98e-11]
    anh = [0.0, 0.0, 0.0, 0.0, 0.0, 1.51e-14]
    for s in det.surf:
        s.n_spec, s.n_diff, s.n_anh = len(spec), len(diff), len(anh)
        for i, v in enumerate(spec):
            s.spec[i] = v
        for i, v in enumerate(diff):
            s.diff[i] = v
        for i, v in enumerate(anh):
            s.anh[i] = v
        s.anh_cutoff, s.refl_cutoff = 520.0, 350.0
    det.max_bounces = max_bounces


def workload(kind: int, mode: int = 0, energy_eV: float = 7.5e-3, origin=(0.0, 0.0, 0.0), logn_median: float = 200.0,
             logn_sigma: float = 1.0, n_min: int = 1, n_max: int = 20000, e_lo_eV: float = 1e-3, e_hi_eV: float = 31e-3,
             pair_at_origin: bool = False, partition: dict | None = None) -> capi.Workload:
    """A device-generated workload. DEPOSIT: a deposit of energy_eV (mode 0 electron
    recoil, 1 nuclear recoil) at `origin` (pair_at_origin set) or uniform in the volume,
    partitioned with the PartitionOptions defaults changed by `partition`."""
    w = capi.Workload()
    w.kind, w.mode, w.energy_eV = kind, mode, energy_eV
    for i in range(3):
        w.origin[i] = origin[i]
    w.logn_median, w.logn_sigma, w.n_min, w.n_max = logn_median, logn_sigma, n_min, n_max
    w.e_lo_eV, w.e_hi_eV, w.pair_at_origin = e_lo_eV, e_hi_eV, int(pair_at_origin)
    w.partition = partition_opts(**(partition or {}))
    return w


# ------------------------------------------------------------------ energy partition
def partition_opts(**changes) -> capi.PartitionOpts:
    """PartitionOptions{} (energy_partition.hpp:22-29) with fields changed; nested
    empirical fields as `empirical_<name>`."""
    o = capi.PartitionOpts()
    check(capi.load().cmpmc_partition_opts_default(C.byref(o)))
    for k, v in changes.items():
        if k.startswith("empirical_"):
            setattr(o.empirical, k[len("empirical_"):], v)
        else:
            setattr(o, k, v)
    return o


def ionization_yield(model: int, energy_eV: float, crystal: capi.Crystal | None = None,
                     empirical: capi.EmpiricalNiel | None = None) -> float:
    """compute_ionization_yield (niel_models.hpp:161-173) for the crystal's target."""
    c = crystal if crystal is not None else germanium()
    return float(capi.load().cmpmc_ionization_yield(model, energy_eV, C.byref(c), C.byref(empirical) if empirical else None))


def lindhard_yield(energy_eV: float, crystal: capi.Crystal | None = None, z1: float = 0.0, a1: float = 0.0) -> float:
    """EnergyPartition::lindhard_ionization_yield (energy_partition.hpp:55-74)."""
    c = crystal if crystal is not None else germanium()
    return float(capi.load().cmpmc_lindhard_yield(energy_eV, C.byref(c), z1, a1))


def sample_fano_binomial(mean: float, fano: float, rng: "Rng") -> int:
    """sample_fano_binomial (niel_models.hpp:203-261) from `rng` (advanced in place)."""
    st = (C.c_uint64 * 4)(*rng.s)
    n = capi.load().cmpmc_sample_fano_binomial(mean, fano, st)
    rng.s = [int(v) for v in st]
    return int(n)


def partition_deposit(lattice: "Lattice", pos, edep_eV: float, nuclear: bool, rng: "Rng",
                      opts: capi.PartitionOpts | None = None, time_s: float = 0.0) -> dict:
    """EnergyPartition::partition_deposit (energy_partition.hpp:78-157) on the host with
    `rng` (advanced in place): {"result", "charges" (CHARGE_DTYPE), "phonons" (PHONON_DTYPE)}."""
    L = capi.load()
    o = opts if opts is not None else partition_opts()
    p = np.ascontiguousarray(pos, dtype=np.float64)
    st = (C.c_uint64 * 4)(*rng.s)
    nq, nph, res = C.c_uint64(), C.c_uint64(), capi.PartitionResult()
    pp = p.ctypes.data_as(C.POINTER(C.c_double))
    # the first call sizes the buffers (a capacity failure leaves the stream untouched)
    rc = L.cmpmc_partition_deposit(C.byref(lattice.c), pp, time_s, edep_eV, int(nuclear), C.byref(o), st, None, 0,
                                   C.byref(nq), None, 0, C.byref(nph), C.byref(res))
    q = np.zeros(max(nq.value, 1), dtype=CHARGE_DTYPE)
    ph = np.zeros(max(nph.value, 1), dtype=PHONON_DTYPE)
    if rc == capi.ERR_CAPACITY:
        rc = L.cmpmc_partition_deposit(C.byref(lattice.c), pp, time_s, edep_eV, int(nuclear), C.byref(o), st,
                                       q.ctypes.data, nq.value, C.byref(nq), ph.ctypes.data, nph.value, C.byref(nph),
                                       C.byref(res))
    check(rc)
    rng.s = [int(v) for v in st]
    return {"result": res.as_dict(), "charges": q[:nq.value].copy(), "phonons": ph[:nph.value].copy()}


class Lattice:
    """CrystalLattice(cfg): builds the octahedral v_g table and the decay inverse CDFs
    on the host (crystal_lattice.hpp:45-55)."""

    def __init__(self, crystal: capi.Crystal | None = None):
        self.crystal = crystal if crystal is not None else germanium()
        self.oct = np.zeros((3, 256, 256, 4), dtype=np.float32)
        self.cdf_lt = np.zeros(1024)
        self.cdf_tt = np.zeros(1024)
        self.c = capi.Lattice()
        check(capi.load().cmpmc_lattice_build(C.byref(self.crystal), self.oct.ctypes.data, self.cdf_lt.ctypes.data,
                                               self.cdf_tt.ctypes.data, C.byref(self.c)))

    KINEMATICS_OCTAHEDRAL, KINEMATICS_EXACT, KINEMATICS_BILINEAR, KINEMATICS_SYMMETRIC = 0, 1, 2, 3

    def build_symmetric(self) -> None:
        """The cubic-symmetry v_g table (kinematics mode 3, cmpmc_lattice_build_symmetric),
        which the batched kernels stage in LDS; engines created afterwards carry it."""
        if getattr(self, "sym", None) is None:
            self.sym = np.zeros((3, 3, 33, 33), dtype=np.float32)
            check(capi.load().cmpmc_lattice_build_symmetric(C.byref(self.c), self.sym.ctypes.data))

    def build_bilinear(self) -> None:
        """The bilinear (cos theta, phi) v_g table (crystal_lattice.hpp:350-368), which
        KinematicsMode::Bilinear reads; engines created afterwards carry it to the device."""
        if getattr(self, "bil", None) is None:
            self.bil = np.zeros((3, 181, 361, 3))
            check(capi.load().cmpmc_lattice_build_bilinear(C.byref(self.c), self.bil.ctypes.data))

    def set_kinematics_mode(self, kin_mode: int) -> None:
        """CrystalLattice::set_kinematics_mode: 0 Octahedral (LUT, default), 1 Exact,
        2 Bilinear (builds the bilinear table), 3 the cubic-symmetry table (this build's;
        builds it). Engines take the lattice's mode when they are created."""
        if kin_mode not in (0, 1, 2, 3):
            raise ValueError("kinematics mode: 0 (Octahedral), 1 (Exact), 2 (Bilinear) or 3 (cubic-symmetry table)")
        if kin_mode == 2:
            self.build_bilinear()
        if kin_mode == 3:
            self.build_symmetric()
        self.c.kinematics_mode = kin_mode

    def map_k_to_vg(self, mode: int, k) -> tuple[np.ndarray, np.ndarray]:
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
        d = np.empty_like(k)
        m = np.empty(k.shape[0])
        check(capi.load().cmpmc_map_k_to_vg(C.byref(self.c), mode, k.shape[0], k.ctypes.data, d.ctypes.data,
                                             m.ctypes.data))
        return d, m


def make_tracks(lattice: Lattice, pos, k, mode, energy, time_s=0.0, weight=1.0) -> np.ndarray:
    """SimulationEngine::make_track for arrays of rows (engine.hpp:171-187)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    k = np.asarray(k, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    mode = np.broadcast_to(np.asarray(mode, dtype=np.int64), (n,))
    nrm2 = (k[:, 0] * k[:, 0] + k[:, 1] * k[:, 1]) + k[:, 2] * k[:, 2]
    inv = np.where(nrm2 > 0, 1.0 / np.sqrt(np.where(nrm2 > 0, nrm2, 1.0)), 0.0)
    ku = np.where(nrm2[:, None] > 0, k * inv[:, None], np.array([0.0, 0.0, 1.0]))
    t = np.zeros(n, dtype=PHONON_DTYPE)
    t["pos"] = pos
    t["k_dir"] = ku
    for m in (0, 1, 2):
        sel = np.nonzero(mode == m)[0]
        if sel.size:
            d, mag = lattice.map_k_to_vg(m, ku[sel])
            t["vg_dir"][sel] = d
            t["vg_mag"][sel] = mag
    t["energy"] = energy
    t["time"] = time_s
    t["weight"] = weight
    t["mode"] = mode
    return t


class Engine:
    """SimulationEngine(lattice, detector) on the current GPU (engine.hpp:158-169)."""

    def __init__(self, lattice: Lattice | None = None, detector: capi.Detector | None = None):
        self.lattice = lattice if lattice is not None else Lattice()
        self.detector = detector if detector is not None else default_detector()
        p = C.c_void_p()
        check(capi.load().cmpmc_model_create(C.byref(self.lattice.c), C.byref(self.detector), C.byref(p)))
        self.ptr = p

    def close(self) -> None:
        if getattr(self, "ptr", None):
            capi.load().cmpmc_model_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cancel(self, value: bool = True) -> None:
        check(capi.load().cmpmc_model_cancel(self.ptr, int(value)))

    def set_kinematics(self, kin_mode: int) -> None:
        """Kinematics of this engine's device model: 0 Octahedral LUT, 1 Exact, 2 Bilinear
        (for an engine whose lattice had that table built), 3 cubic-symmetry table (every
        model carries it)."""
        check(capi.load().cmpmc_model_set_kinematics(self.ptr, kin_mode))

    def device_map_k_to_vg(self, mode: int, k, form: int = FORM_PLAIN) -> tuple[np.ndarray, np.ndarray]:
        """map_k_to_vg as the kernels of `form` compute it (FP64 plain, FP32 batched)."""
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
        d = np.empty_like(k)
        m = np.empty(k.shape[0])
        check(capi.load().cmpmc_model_map_k_to_vg(self.ptr, mode, form, k.shape[0], k.ctypes.data, d.ctypes.data,
                                                  m.ctypes.data, None))
        return d, m

    def field_at(self, pts, form: int = FORM_PLAIN, start=None) -> dict:
        """The electric field as the charge step evaluates it on the device, at 3D points:
        {"inside", "cell", "E"}; with a mesh each walk starts at start[i] (None: the cell a
        new carrier starts from)."""
        p = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = len(p)
        st = None if start is None else np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.int32), (n,)))
        inside, cell, E = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros((n, 3))
        check(capi.load().cmpmc_model_field_at(self.ptr, form, n, p.ctypes.data, None if st is None else st.ctypes.data,
                                               inside.ctypes.data, cell.ctypes.data, E.ctypes.data, None))
        return {"inside": inside, "cell": cell, "E": E}

    def trim(self) -> None:
        check(capi.load().cmpmc_model_trim(self.ptr))

    @staticmethod
    def _opts(form, seed, keep_hits, first_branch, block, headroom, lane_depth, grid_blocks, refill_period,
              poll_cancel, waves=0, time_budget_s=0.0) -> capi.RunOpts:
        o = capi.RunOpts()
        o.form, o.keep_hits, o.want_first_branch, o.block_size = form, int(keep_hits), int(first_branch), block
        o.seed, o.stack_headroom, o.lane_stack_depth = seed, headroom, lane_depth
        o.grid_blocks, o.refill_period, o.poll_cancel = grid_blocks, refill_period, int(poll_cancel)
        o.waves_per_simd = waves
        o.time_budget_s = time_budget_s
        return o

    def hit_estimate(self, wl: capi.Workload, n_events: int) -> int:
        """A generous number of sensor hits for n_events of `wl` (the default hit
        capacity, so that a run rarely has to be repeated for room): a phonon of energy
        E ends in about E / 0.3 meV film hits at the default films; a pair yields its two
        collected carriers and, when Luke phonons are tracked, about 4 eV of them. Capped
        at 2^27 hits (7.5 GB of host records): a larger run that keeps its hits sizes the
        buffer exactly on a second pass."""
        e = 0.5 * (wl.e_lo_eV + wl.e_hi_eV) if wl.kind == WL_PHONON_LOGN else wl.energy_eV
        per_phonon = 1.5 * max(e, 0.0) / 3.0e-4 + 8
        if wl.kind == WL_PHONON_POINT:
            n = per_phonon
        elif wl.kind == WL_PHONON_UNIF:
            n = per_phonon * 0.5 * (wl.n_min + wl.n_max)
        elif wl.kind == WL_PHONON_LOGN:
            mean = wl.logn_median * math.exp(0.5 * wl.logn_sigma ** 2)
            n = per_phonon * min(mean, float(wl.n_max))
        elif wl.kind == WL_DEPOSIT:
            # prompt phonons near the Debye energy (~36 meV in Ge), both carriers of each pair
            # and, when Luke phonons are tracked, about 4 eV of them per pair
            pairs = min(float(wl.partition.max_sampled_pairs), wl.energy_eV / 2.96 + 1.0)
            phon = min(float(wl.partition.max_sampled_phonons), wl.energy_eV / 0.036 + 1.0)
            luke = 1.5 * 4.0 / 3.0e-4 * min(max(self.detector.luke_sample_fraction, 0.0), 1.0)
            n = 2 * pairs * (1.0 + luke) + phon * (1.5 * 0.036 / 3.0e-4 + 8)
        else:
            n = 4 + (1.5 * 4.0 / 3.0e-4 * min(max(self.detector.luke_sample_fraction, 0.0), 1.0))
        return int(min(max(1024, n * n_events), 1 << 27))

    def _run(self, call, n_events: int, keep_hits: bool, hit_capacity: int | None, first_branch: bool) -> dict:
        cap = hit_capacity if hit_capacity is not None else (max(1024, 64 * n_events) if keep_hits else 0)
        for _ in range(3):
            out = capi.RunOut()
            hits = np.zeros(max(cap, 1), dtype=HIT_DTYPE)
            offs = np.zeros(n_events + 1, dtype=np.uint64)
            fb = np.zeros(max(n_events, 1), dtype=np.int32)
            out.hits = hits.ctypes.data if keep_hits else None
            out.hit_capacity = cap if keep_hits else 0
            out.hit_offsets = offs.ctypes.data_as(C.POINTER(capi.u64))
            out.first_branch = fb.ctypes.data_as(C.POINTER(capi.i32)) if first_branch else None
            t0 = time.perf_counter()
            rc = call(out)
            wall = time.perf_counter() - t0
            if rc == capi.ERR_CAPACITY and keep_hits and out.hits_needed > cap:
                cap = int(out.hits_needed)      # rerun with the room the library asks for
                continue
            check(rc)
            res = {"stats": out.stats.as_dict(), "stats_raw": bytes(out.stats), "hits": hits[:out.n_hits].copy(),
                   "hit_offsets": offs, "kernel_ms": out.kernel_ms, "wall_s": wall, "passes": out.passes}
            if first_branch:
                res["first_branch"] = fb[:n_events].copy()
            return res
        raise CmpmcError(capi.ERR_CAPACITY, "hit capacity retries exhausted")

    def run_workload(self, wl: capi.Workload, n_events: int, seed: int = 1, form: int = FORM_PLAIN,
                     keep_hits: bool = True, hit_capacity: int | None = None, first_branch: bool = False,
                     block: int = 0, headroom: int = 0, lane_depth: int = 0, grid_blocks: int = 0,
                     refill_period: int = 0, poll_cancel: bool = False, waves: int = 0, time_budget_s: float = 0.0,
                     stream=None, no_alloc: bool = False) -> dict:
        """run_events with the generator on the device (parallel.hpp:76-148). time_budget_s > 0
        stops the run with CMPMC_ERR_CANCELLED once it has run that long. no_alloc: run
        through cmpmc_run_workload_noalloc (every device buffer from the library's fixed
        per-device scratch; no device allocation)."""
        o = self._opts(form, seed, keep_hits, first_branch, block, headroom, lane_depth, grid_blocks, refill_period,
                       poll_cancel, waves, time_budget_s)
        L = capi.load()
        fn = L.cmpmc_run_workload_noalloc if no_alloc else L.cmpmc_run_workload
        if keep_hits and hit_capacity is None:
            hit_capacity = self.hit_estimate(wl, n_events)
        return self._run(lambda out: fn(self.ptr, C.byref(wl), n_events, C.byref(o), C.byref(out), stream),
                         n_events, keep_hits, hit_capacity, first_branch)

    def autotune(self, wl: capi.Workload, n_events: int, seed: int = 1, cache: str | None = None,
                 force: bool = False, stream=None) -> dict:
        """Launch options of the batched form for `wl` on this device (cmpmc_autotune):
        coordinate descent over timed runs of n_events events, or the entry of the cache
        file. Returns run_workload / plan keyword arguments (block, waves, refill_period,
        lane_depth) and the timings (best_ms, default_ms, trials, from_cache)."""
        o = self._opts(FORM_BATCHED, seed, False, False, 0, 0, 0, 0, 0, False, 0)
        r = capi.TuneResult()
        check(capi.load().cmpmc_autotune(self.ptr, C.byref(wl), n_events, seed,
                                         cache.encode() if cache else None, capi.TUNE_FORCE if force else 0,
                                         C.byref(o), C.byref(r), stream))
        return {"block": o.block_size, "waves": o.waves_per_simd, "refill_period": o.refill_period,
                "lane_depth": o.lane_stack_depth, "best_ms": r.best_ms, "default_ms": r.default_ms,
                "trials": r.trials, "from_cache": bool(r.from_cache)}

    def plan(self, wl: capi.Workload, n_events: int, keep_hits: bool = True, hit_capacity: int | None = None,
             block: int = 0, lane_depth: int = 0, grid_blocks: int = 0, refill_period: int = 0,
             poll_cancel: bool = False, waves: int = 0, stream=None) -> "Plan":
        """A graph-captured batched run of n_events events of `wl` (cmpmc_plan_create);
        Plan.run(seed, first_event) replays it."""
        return Plan(self, wl, n_events, keep_hits, hit_capacity,
                    self._opts(FORM_BATCHED, 1, keep_hits, False, block, 0, lane_depth, grid_blocks, refill_period,
                               poll_cancel, waves), stream)

    def run_events(self, tracks: np.ndarray, track_offsets: np.ndarray, rng_state: np.ndarray,
                   injected: np.ndarray | None = None, charges: np.ndarray | None = None,
                   charge_offsets: np.ndarray | None = None, seed: int = 1, form: int = FORM_PLAIN,
                   keep_hits: bool = True, hit_capacity: int | None = None, first_branch: bool = False,
                   first_event_id: int = 0, **kw) -> dict:
        """run_events for stacks produced by a host generator: tracks of event i are rows
        track_offsets[i]:track_offsets[i+1]; rng_state[i] is its post-generator stream."""
        n = len(track_offsets) - 1
        tracks = np.ascontiguousarray(tracks, dtype=PHONON_DTYPE)
        toff = np.ascontiguousarray(track_offsets, dtype=np.uint64)
        rng = np.ascontiguousarray(rng_state, dtype=np.uint64).reshape(n, 4)
        b = capi.EventBatch()
        b.n_events = n
        b.tracks = tracks.ctypes.data if tracks.size else None
        b.track_offsets = toff.ctypes.data_as(C.POINTER(capi.u64))
        keep = [tracks, toff, rng]
        if charges is not None:
            charges = np.ascontiguousarray(charges, dtype=CHARGE_DTYPE)
            coff = np.ascontiguousarray(charge_offsets, dtype=np.uint64)
            b.charges = charges.ctypes.data
            b.charge_offsets = coff.ctypes.data_as(C.POINTER(capi.u64))
            keep += [charges, coff]
        b.rng_state = rng.ctypes.data_as(C.POINTER(capi.u64))
        if injected is not None:
            inj = np.ascontiguousarray(injected, dtype=np.float64)
            b.injected_eV = inj.ctypes.data_as(C.POINTER(capi.dbl))
            keep.append(inj)
        b.first_event_id = first_event_id
        o = self._opts(form, seed, keep_hits, first_branch, kw.get("block", 0), kw.get("headroom", 0),
                       kw.get("lane_depth", 0), kw.get("grid_blocks", 0), kw.get("refill_period", 0),
                       kw.get("poll_cancel", False), kw.get("waves", 0))
        L = capi.load()
        return self._run(lambda out: L.cmpmc_run_events(self.ptr, C.byref(b), C.byref(o), C.byref(out), None),
                         n, keep_hits, hit_capacity, first_branch)

    def diff_forward(self, wl: capi.Workload, n_events: int, params, seed: int = 1, t_max_s: float = 1e-3,
                     waves: int = 0, grid_blocks: int = 0, stream=None) -> tuple[np.ndarray, "DiffTape", dict]:
        """Differentiable batched transport of a phonon workload: observables y (top and
        bottom face collected energy, 64-bin energy-weighted arrival-time histogram over
        [0, t_max_s)), a tape for the backward pass and the run's statistics.
        params = (B_iso, A_anh, f_diffuse, p_film)."""
        p = capi.DiffParams(*[float(v) for v in params])
        o = self._opts(FORM_BATCHED, seed, False, False, 0, 0, 0, grid_blocks, 0, False, waves)
        y = np.zeros(capi.DIFF_OUT)
        st = capi.Stats()
        tape = C.c_void_p()
        ms = C.c_double()
        t0 = time.perf_counter()
        check(capi.load().cmpmc_diff_forward(self.ptr, C.byref(wl), n_events, C.byref(p), t_max_s, C.byref(o),
                                             y.ctypes.data_as(C.POINTER(capi.dbl)), C.byref(st), C.byref(tape),
                                             C.byref(ms), stream))
        info = {"stats": st.as_dict(), "kernel_ms": ms.value, "wall_s": time.perf_counter() - t0}
        return y, DiffTape(tape, self), info

    def transport(self, tracks: np.ndarray, rng_state: np.ndarray, hit_capacity: int = 1 << 16) -> dict:
        """SimulationEngine::transport(stack, hits, stats, rng) (engine.hpp:403-582)."""
        tracks = np.ascontiguousarray(tracks, dtype=PHONON_DTYPE)
        rng = np.ascontiguousarray(rng_state, dtype=np.uint64).copy()
        L = capi.load()
        while True:
            hits = np.zeros(hit_capacity, dtype=HIT_DTYPE)
            nh = capi.u64(0)
            st = capi.Stats()
            r = rng.copy()
            rc = L.cmpmc_transport_serial(self.ptr, tracks.ctypes.data, len(tracks),
                                          r.ctypes.data_as(C.POINTER(capi.u64)), hits.ctypes.data, hit_capacity,
                                          C.byref(nh), C.byref(st), None)
            if rc == capi.ERR_CAPACITY and hit_capacity < (1 << 26):
                hit_capacity *= 8
                continue
            check(rc)
            return {"stats": st.as_dict(), "stats_raw": bytes(st), "hits": hits[:nh.value].copy(), "rng": r}

    def transport_charges(self, charges: np.ndarray, rng_state: np.ndarray, phonon_capacity: int = 1 << 16,
                          hit_capacity: int = 1 << 12) -> dict:
        """SimulationEngine::transport_charges (engine.hpp:192-380): returns the hits and the
        Luke phonons it pushed (to be transported by a following transport call)."""
        charges = np.ascontiguousarray(charges, dtype=CHARGE_DTYPE)
        rng = np.ascontiguousarray(rng_state, dtype=np.uint64).copy()
        L = capi.load()
        while True:
            ph = np.zeros(phonon_capacity, dtype=PHONON_DTYPE)
            hits = np.zeros(hit_capacity, dtype=HIT_DTYPE)
            nph, nh = capi.u64(0), capi.u64(0)
            st = capi.Stats()
            r = rng.copy()
            rc = L.cmpmc_transport_charges_serial(self.ptr, charges.ctypes.data, len(charges),
                                                  r.ctypes.data_as(C.POINTER(capi.u64)), ph.ctypes.data,
                                                  phonon_capacity, C.byref(nph), hits.ctypes.data, hit_capacity,
                                                  C.byref(nh), C.byref(st), None)
            if rc == capi.ERR_CAPACITY and phonon_capacity < (1 << 24):
                phonon_capacity *= 8
                hit_capacity *= 8
                continue
            check(rc)
            return {"stats": st.as_dict(), "stats_raw": bytes(st), "hits": hits[:nh.value].copy(),
                    "phonons": ph[:nph.value].copy(), "rng": r}


class DiffTape:
    """Hits of a differentiable forward pass with their scores (device memory), consumed
    by backward(); freed with the object."""

    def __init__(self, ptr: C.c_void_p, engine: Engine):
        self.ptr = ptr
        self.engine = engine          # the model outlives its tapes
        self.last_kernel_ms = 0.0

    def __len__(self) -> int:
        return int(capi.load().cmpmc_diff_tape_size(self.ptr)) if self.ptr else 0

    def backward(self, g, stream=None) -> np.ndarray:
        """grad[j] = sum_k g[k] dE[y_k]/dtheta_j for theta = (B, A, f_diffuse, p_film)."""
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(capi.DIFF_OUT)
        out = np.zeros(4)
        ms = C.c_double()
        check(capi.load().cmpmc_diff_backward(self.ptr, g.ctypes.data_as(C.POINTER(capi.dbl)),
                                              out.ctypes.data_as(C.POINTER(capi.dbl)), C.byref(ms), stream))
        self.last_kernel_ms = ms.value
        return out

    def close(self) -> None:
        if getattr(self, "ptr", None):
            capi.load().cmpmc_diff_tape_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A batched run of a fixed device workload recorded as a HIP graph (cmpmc_plan_*).
    run(seed, first_event) replays it and returns what Engine.run_workload returns for
    those events and seed; buffers that prove too small are enlarged and the graph
    recorded again (`recordings`)."""

    def __init__(self, engine: Engine, wl: capi.Workload, n_events: int, keep_hits: bool,
                 hit_capacity: int | None, opts: capi.RunOpts, stream=None):
        self.engine = engine          # the model outlives its plans
        self.n_events = n_events
        self.keep_hits = keep_hits
        self.hit_capacity = hit_capacity if hit_capacity is not None else (max(1024, 64 * n_events) if keep_hits else 0)
        p = C.c_void_p()
        check(capi.load().cmpmc_plan_create(engine.ptr, C.byref(wl), n_events, C.byref(opts),
                                            self.hit_capacity if keep_hits else 0, stream, C.byref(p)))
        self.ptr = p

    @property
    def recordings(self) -> int:
        return int(capi.load().cmpmc_plan_recordings(self.ptr))

    def run(self, seed: int, first_event: int = 0, stream=None) -> dict:
        L = capi.load()
        return self.engine._run(lambda out: L.cmpmc_plan_run(self.ptr, seed, first_event, C.byref(out), stream),
                                self.n_events, self.keep_hits, self.hit_capacity, False)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            capi.load().cmpmc_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
