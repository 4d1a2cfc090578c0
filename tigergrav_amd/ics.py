"""Seeded synthetic particle sets for the benchmark configurations.

All generators return int32 [N, 3] fixed-point positions (x = p 2^-32 + 0.5, the reference's
pos_type, tigergrav/position.hpp:27-51).

  uniform(N, seed)                      i.i.d. U[0,1)^3
  plummer_spheres(N, seed, ...)         8 Plummer spheres (scale radius a, centres U[0,1)^3, equal
                                        counts) + a uniform background fraction; samples outside
                                        [0,1)^3 are redrawn (open boundaries)
  zeldovich(ngrid, seed, n_s, rms)      ngrid^3 lattice displaced by the Zel'dovich field
                                        psi_k = i k delta_k / k^2 of a Gaussian random field with
                                        P(k) ~ k^n_s, zero beyond the grid Nyquist |k| > pi ngrid,
                                        scaled to rms |psi| = rms * cell; positions wrapped mod 1
"""
import numpy as np


def double_to_pos(d):
    """lround((d - 0.5) 2^32) wrapped into int32 (tigergrav/position.hpp:44-51)."""
    d = np.asarray(d, np.float64)
    y = (d - 0.5) * 4294967296.0
    z = np.where(y >= 0, np.floor(y + 0.5), np.ceil(y - 0.5)).astype(np.int64)  # half away from zero
    z = np.where(z >= 2 ** 31, z - 2 ** 32, z)
    return z.astype(np.int32)


def pos_to_double(p):
    return np.asarray(p, np.float64) * (1.0 / 4294967296.0) + 0.5


def uniform(n, seed=1):
    rng = np.random.default_rng(seed)
    return double_to_pos(rng.random((n, 3)))


def plummer_spheres(n, seed=2, n_spheres=8, a=0.02, background=0.2):
    rng = np.random.default_rng(seed)
    n_sph = int(n * (1.0 - background)) // n_spheres
    centres = rng.random((n_spheres, 3))
    out = []
    for c in centres:
        pts = np.empty((0, 3))
        while len(pts) < n_sph:
            k = n_sph - len(pts)
            u = rng.random(k)
            u = np.clip(u, 1e-12, 1.0)
            r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
            v = rng.normal(size=(k, 3))
            v /= np.linalg.norm(v, axis=1)[:, None]
            p = c[None, :] + r[:, None] * v
            ok = np.all((p >= 0.0) & (p < 1.0), axis=1)
            pts = np.concatenate([pts, p[ok]])
        out.append(pts[:n_sph])
    nb = n - n_spheres * n_sph
    out.append(rng.random((nb, 3)))
    x = np.concatenate(out)
    return double_to_pos(x)


def zeldovich(ngrid, seed=3, n_s=-2.0, rms_cells=0.3, return_psi=False):
    rng = np.random.default_rng(seed)
    ng = int(ngrid)
    white = rng.normal(size=(ng, ng, ng))
    dk = np.fft.rfftn(white)
    kx = 2.0 * np.pi * np.fft.fftfreq(ng, d=1.0 / ng)
    kz = 2.0 * np.pi * np.fft.rfftfreq(ng, d=1.0 / ng)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, k ** (n_s / 2.0), 0.0)
    amp[k > np.pi * ng] = 0.0
    dk *= amp
    psi = np.empty((ng, ng, ng, 3))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(k2 > 0, 1.0 / k2, 0.0)
    for d, K in enumerate((KX, KY, KZ)):
        psi[..., d] = np.fft.irfftn(1j * K * inv * dk, s=(ng, ng, ng), axes=(0, 1, 2))
    rms = np.sqrt(np.mean(np.sum(psi * psi, axis=-1)))
    psi *= (rms_cells / ng) / rms
    q = (np.arange(ng) + 0.5) / ng
    Q = np.stack(np.meshgrid(q, q, q, indexing="ij"), axis=-1)
    x = np.mod(Q + psi, 1.0).reshape(-1, 3)
    pos = double_to_pos(x)
    if return_psi:
        return pos, psi.reshape(-1, 3)
    return pos


def soft_len(n):
    """The reference's default softening length 0.02 N^(-1/3) (src/options.cpp:86-88)."""
    return 0.02 * float(n) ** (-1.0 / 3.0)


def load_ngenic(path):
    """An N-GenIC / Gadget-1 file (include/tg_load.h, src/load.cpp:34-96): (x int32 [N, 3],
    v float32 [N, 3], header dict).  Positions shifted by half the mean spacing and wrapped,
    velocities times (1 + z)^(-3/2), as the reference loads them."""
    import ctypes as C

    from ._lib import TGError, lib

    class Header(C.Structure):
        _fields_ = [("npart", C.c_uint32 * 6), ("mass", C.c_double * 6), ("time", C.c_double), ("redshift", C.c_double),
                    ("flag_sfr", C.c_int32), ("flag_feedback", C.c_int32), ("npartTotal", C.c_uint32 * 6),
                    ("flag_cooling", C.c_int32), ("num_files", C.c_int32), ("BoxSize", C.c_double), ("Omega0", C.c_double),
                    ("OmegaLambda", C.c_double), ("HubbleParam", C.c_double), ("flag_stellarage", C.c_int32),
                    ("flag_metals", C.c_int32), ("hashtabsize", C.c_int32), ("fill", C.c_char * 84)]

    assert C.sizeof(Header) == 256
    L = lib()
    L.tg_ngenic_read_header.argtypes = [C.c_char_p, C.POINTER(Header)]
    L.tg_ngenic_load.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    h = Header()
    p = str(path).encode()
    if L.tg_ngenic_read_header(p, C.byref(h)) != 0:
        raise TGError(L.tg_last_error().decode(errors="replace"))
    n = C.c_int64()
    npart = int(h.npart[1])
    x = np.zeros((npart, 3), np.int32)
    v = np.zeros((npart, 3), np.float32)
    if L.tg_ngenic_load(p, npart, x.ctypes.data, v.ctypes.data, C.byref(n)) != 0:
        raise TGError(L.tg_last_error().decode(errors="replace"))
    hd = {name: (list(getattr(h, name)) if name in ("npart", "mass", "npartTotal") else getattr(h, name))
          for name, _ in Header._fields_ if name != "fill"}
    return x, v, hd
