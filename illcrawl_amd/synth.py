"""Synthetic gas-cell snapshots and cameras for the projective smoothing path.

No snapshot files exist offline, so every workload is generated from a seed:

* positions: a fraction ``frac_uniform`` uniform in the periodic box, the rest
  split evenly over NFW halos (radius by inverse CDF of
  ``mu(x) = ln(1+x) - x/(1+x)`` on ``x in [0, c]``, isotropic directions),
  wrapped into ``[0, L)``;
* masses: float32 lognormal with mean ``mass_mean`` and log-sigma ``sigma``;
* density: 32-nearest-neighbour estimate
  ``rho_i = sum_{j in 32NN(i)} m_j / (4/3 pi r_32^3)`` (periodic);
* smoothing length ``h = 2.5 (m/rho)^(1/3)``;
* gas attributes (X-ray workloads): temperature log-uniform in
  ``[1e5, 1e8]`` K, multiplied by 3 inside ``r_s`` of any halo; electron
  abundance uniform in ``[1.0, 1.2]``.

The camera helpers build the view that illcrawl's
``create_distribution_centered_camera`` builds (camera at the centre minus
``R * look_at``, integration depth ``2R``; illcrawl_app.cpp:179-204), with
``R`` the half-diagonal of the box.
"""
from __future__ import annotations

import dataclasses
import math
import sys

import numpy as np


@dataclasses.dataclass
class Snapshot:
    box: float
    pos: np.ndarray          # (N, 3) float32
    mass: np.ndarray         # (N,) float32
    rho: np.ndarray          # (N,) float32
    hsml: np.ndarray         # (N,) float32
    temp: np.ndarray | None = None   # (N,) float32, K
    xe: np.ndarray | None = None     # (N,) float32
    halos: list | None = None

    @property
    def n(self):
        return self.pos.shape[0]

    def particles4(self, w=None):
        """float4 (x, y, z, w) array, the reference's particle layout
        (particle_grid.hpp:43-48); w defaults to the mass."""
        p4 = np.empty((self.n, 4), np.float32)
        p4[:, :3] = self.pos
        p4[:, 3] = self.mass if w is None else w
        return p4


def _nfw_inverse_cdf(u, c):
    """x with mu(x) = u * mu(c), mu(x) = ln(1+x) - x/(1+x); closed form via
    the principal Lambert W branch."""
    from scipy.special import lambertw
    muc = math.log1p(c) - c / (1.0 + c)
    p = u * muc
    w = lambertw(-np.exp(-1.0 - p), 0).real
    x = -1.0 / w - 1.0
    return np.clip(x, 0.0, c)


def knn_density(pos, mass, box, k=32, workers=-1):
    """rho_i = (sum of the k neighbours' masses) / (4/3 pi r_k^3), periodic."""
    from scipy.spatial import cKDTree
    p = np.asarray(pos, np.float64) % box
    tree = cKDTree(p, boxsize=box, leafsize=32, balanced_tree=False, compact_nodes=False)
    m64 = np.asarray(mass, np.float64)
    rho = np.empty(p.shape[0], np.float32)
    chunk = 1 << 21   # bounded memory for large snapshots; queries are independent
    for a in range(0, p.shape[0], chunk):
        if p.shape[0] > (1 << 23):
            print(f"[synth] density {a}/{p.shape[0]}", file=sys.stderr, flush=True)
        d, idx = tree.query(p[a:a + chunk], k=k + 1, workers=workers)
        rk = d[:, k]
        msum = m64[idx[:, 1:]].sum(axis=1)
        rho[a:a + chunk] = (msum / (4.0 / 3.0 * math.pi * rk ** 3)).astype(np.float32)
    return rho


def make_snapshot(n, seed, box, halos=None, n_halos=1, rs_range=(300.0, 300.0),
                  conc=5.0, frac_uniform=0.5, mass_mean=1e-4, sigma=0.3,
                  gas=False, knn=32):
    """Generate a seeded snapshot.  ``halos`` is a list of (centre, r_s);
    when None, ``n_halos`` halos are drawn: the first at the box centre, the
    rest uniform in the central half of the box, r_s uniform in rs_range."""
    rng = np.random.default_rng(seed)
    if halos is None:
        halos = []
        for k in range(n_halos):
            if k == 0:
                c = np.full(3, 0.5 * box)
            else:
                c = rng.uniform(0.25 * box, 0.75 * box, 3)
            rs = rng.uniform(rs_range[0], rs_range[1]) if rs_range[1] > rs_range[0] \
                else rs_range[0]
            halos.append((np.asarray(c, np.float64), float(rs)))
    n_uni = int(round(n * frac_uniform))
    n_halo_tot = n - n_uni
    pos = np.empty((n, 3), np.float64)
    pos[:n_uni] = rng.uniform(0.0, box, (n_uni, 3))
    halo_id = np.full(n, -1, np.int32)
    start = n_uni
    for k, (c, rs) in enumerate(halos):
        cnt = n_halo_tot // len(halos) + (1 if k < n_halo_tot % len(halos) else 0)
        x = _nfw_inverse_cdf(rng.random(cnt), conc)
        r = x * rs
        cost = rng.uniform(-1.0, 1.0, cnt)
        phi = rng.uniform(0.0, 2.0 * math.pi, cnt)
        sint = np.sqrt(1.0 - cost ** 2)
        d = np.stack([sint * np.cos(phi), sint * np.sin(phi), cost], axis=1)
        pos[start:start + cnt] = c + r[:, None] * d
        halo_id[start:start + cnt] = k
        start += cnt
    pos %= box
    mu = math.log(mass_mean) - 0.5 * sigma * sigma
    mass = np.exp(rng.normal(mu, sigma, n)).astype(np.float32)
    pos32 = pos.astype(np.float32)
    rho = knn_density(pos32, mass, box, k=knn)
    hsml = (2.5 * np.cbrt(mass.astype(np.float64) / rho.astype(np.float64))).astype(np.float32)
    snap = Snapshot(box=float(box), pos=pos32, mass=mass, rho=rho, hsml=hsml,
                    halos=[(np.asarray(c), rs) for c, rs in halos])
    if gas:
        temp = np.exp(rng.uniform(math.log(1e5), math.log(1e8), n))
        inside = np.zeros(n, bool)
        for c, rs in halos:
            d = pos - c
            d -= box * np.round(d / box)
            inside |= (d * d).sum(axis=1) < rs * rs
        temp[inside] *= 3.0
        snap.temp = temp.astype(np.float32)
        snap.xe = rng.uniform(1.0, 1.2, n).astype(np.float32)
    return snap


# ------------------------------------------------------------ workloads

def config1(n=1 << 20, seed=0):
    """Plain projected gas density: 2^20 cells, seed 0, 10000 kpc/h box,
    one NFW halo (r_s=300, c=5) at the centre."""
    return make_snapshot(n, seed, 10000.0, n_halos=1, rs_range=(300.0, 300.0))


def config2(n=1 << 24, seed=1, n_halos=8):
    """X-ray workload: 16M cells, seed 1, 75000 kpc/h box, 8 NFW halos with
    r_s uniform in [100, 500]."""
    return make_snapshot(n, seed, 75000.0, n_halos=n_halos, rs_range=(100.0, 500.0), gas=True)


def config4(n=1 << 22, seed=3):
    """Adjoint workload: 4M cells of the config-1 generator at seed 3."""
    return make_snapshot(n, seed, 10000.0, n_halos=1, rs_range=(300.0, 300.0))


def config5(n=1 << 26, seed=4, n_halos=32):
    """Slab tomography workload: 64M cells of the X-ray generator at seed 4 with
    32 halos in the 75000 kpc/h box."""
    return make_snapshot(n, seed, 75000.0, n_halos=n_halos, rs_range=(100.0, 500.0), gas=True)


def slab_views(box, n_slabs=32, thickness=500.0, nx=4096, width=None, quantity=1):
    """Tomography slabs along +z through the box centre: slab k is a view of
    depth ``thickness`` whose screen sits at z0 + k * thickness (z0 = centre -
    n_slabs * thickness / 2); the depth cut is the reference's centre mask."""
    c = 0.5 * box
    z0 = c - 0.5 * n_slabs * thickness
    width = box if width is None else width
    views = []
    for k in range(n_slabs):
        views.append(dict(position=np.array([c, c, z0 + k * thickness]),
                          look_at=np.array([0.0, 0.0, 1.0]), roll=0.0, width=float(width),
                          nx=int(nx), ny=int(nx), depth=float(thickness), quantity=quantity))
    return views


def centred_view(box, look_at=(0.0, 0.0, 1.0), width=2000.0, nx=512, ny=None, roll=0.0,
                 centre=None):
    """Camera dict for a view through the box centre: position = centre -
    R look_at, depth 2R, R = half box diagonal (illcrawl_app.cpp:179-204)."""
    la = np.asarray(look_at, np.float64)
    la = la / np.linalg.norm(la)
    c = np.full(3, 0.5 * box) if centre is None else np.asarray(centre, np.float64)
    R = 0.5 * math.sqrt(3.0) * box
    return dict(position=c - R * la, look_at=la, roll=float(roll), width=float(width),
                nx=int(nx), ny=int(ny if ny is not None else nx), depth=2.0 * R)


def random_views(box, n_views=64, seed=2, resolutions=(256, 512, 1024, 2048),
                 widths=(1000.0, 5000.0)):
    """The batched multi-view workload: random unit look_at and roll from
    ``seed``, resolutions in equal groups, widths uniform, quantity alternating
    column density (0) / X-ray (1)."""
    rng = np.random.default_rng(seed)
    views = []
    per = n_views // len(resolutions)
    for k in range(n_views):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        roll = rng.uniform(0.0, 2.0 * math.pi)
        width = rng.uniform(widths[0], widths[1])
        res = resolutions[min(k // per, len(resolutions) - 1)]
        d = centred_view(box, look_at=v, width=width, nx=res, roll=roll)
        d["quantity"] = k % 2
        views.append(d)
    return views


def gaunt_table_1d(n=512, t_min=1e4, t_max=1e9):
    """512-entry tabulated frequency-averaged free-free Gaunt factor on ln T,
    generated from the analytic fit g(T) = 1.1 + 0.34 exp(-(5.5 - log10 T)^2 / 3)
    (Rybicki & Lightman fig. 5.2 fit).  Returns (y float32[n], x0 = ln t_min,
    dx) with node i at x0 + i*dx."""
    x0 = math.log(t_min)
    dx = (math.log(t_max) - x0) / (n - 1)
    lnT = x0 + dx * np.arange(n)
    lg = lnT / math.log(10.0)
    y = 1.1 + 0.34 * np.exp(-((5.5 - lg) ** 2) / 3.0)
    return y.astype(np.float32), np.float32(x0), np.float32(dx)


def cached(name, cache_dir=None):
    """Workload snapshot by name (``c1``, ``c2``, ``c4``, ``c5``), cached as an .npz
    under ``cache_dir`` (default $ILC_CACHE or /tmp/ilc_cache) so that one
    session generates it once.  Loaded with numpy's pickle-free loader."""
    import os
    cache_dir = cache_dir or os.environ.get("ILC_CACHE", "/tmp/ilc_cache")
    path = os.path.join(cache_dir, f"{name}.npz")
    makers = {"c1": config1, "c2": config2, "c4": config4, "c5": config5}
    if os.path.exists(path):
        z = np.load(path)
        halos = [(z["halo_c"][i], float(z["halo_rs"][i])) for i in range(z["halo_c"].shape[0])]
        return Snapshot(box=float(z["box"]), pos=z["pos"], mass=z["mass"], rho=z["rho"],
                        hsml=z["hsml"], temp=z["temp"] if "temp" in z else None,
                        xe=z["xe"] if "xe" in z else None, halos=halos)
    snap = makers[name]()
    os.makedirs(cache_dir, exist_ok=True)
    extra = {}
    if snap.temp is not None:
        extra = dict(temp=snap.temp, xe=snap.xe)
    tmp = path + ".tmp.npz"
    np.savez(tmp, box=snap.box, pos=snap.pos, mass=snap.mass, rho=snap.rho, hsml=snap.hsml,
             halo_c=np.array([c for c, _ in snap.halos]),
             halo_rs=np.array([rs for _, rs in snap.halos]), **extra)
    os.replace(tmp, path)
    return snap
