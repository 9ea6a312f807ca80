"""Synthetic molecular inputs for the solver (input recipes; nothing here runs on the GPU).

Atoms are rows ``(i, j, k, Z, rc)`` in fractional grid units, the layout of
``cloud/solver.py:378-382``; the config-dict form ``{'i','j','k','Z','rc','el'}``
is what ``RealQMSolver`` takes (``cloud/run_h2o_test.py:32-39``).

Element conventions follow the reference builders: O carries Z = 2 valence
electrons with rc = 0.4 bohr, H Z = 1 with rc = 0, C Z = 4 rc = 0.3, N Z = 3
rc = 0.3 (``cloud/run_h2o_test.py:35-37``; ``cloud/run_hairpin.py:38``).
"""
from __future__ import annotations

import math

import numpy as np

ANGSTROM = 1.8897  # bohr per angstrom, cloud/run_hairpin.py:29
ELEMENT = {"H": (1, 0.0), "O": (2, 0.4), "C": (4, 0.3), "N": (3, 0.3)}


def to_config_atoms(rows, els=None):
    out = []
    for n, r in enumerate(rows):
        d = {"i": float(r[0]), "j": float(r[1]), "k": float(r[2]), "Z": float(r[3]), "rc": float(r[4])}
        if els is not None:
            d["el"] = els[n]
        out.append(d)
    return out


def h2o_reference_builder(NN=100, screen=10.0, rOH=1.81, angle_deg=104.5):
    """The reference H2O builder with integer positions (cloud/run_h2o_test.py:24-39)."""
    h = screen / NN
    N2 = NN // 2
    ang = angle_deg * math.pi / 180
    Oi, Oj = N2, N2
    H1i = round(Oi + rOH * math.cos(math.pi / 2 + ang / 2) / h)
    H1j = round(Oj - rOH * math.sin(math.pi / 2 + ang / 2) / h)
    H2i = round(Oi + rOH * math.cos(math.pi / 2 - ang / 2) / h)
    H2j = round(Oj - rOH * math.sin(math.pi / 2 - ang / 2) / h)
    atoms = [
        {"i": H1i, "j": H1j, "Z": 1, "el": "H", "rc": 0},
        {"i": Oi, "j": Oj, "Z": 2, "el": "O", "rc": 0.4},
        {"i": H2i, "j": H2j, "Z": 1, "el": "H", "rc": 0},
    ]
    return {"gridN": NN, "screen": screen, "atoms": atoms}


def h2o_single(NN=96, h=0.1, rOH=1.809, angle_deg=104.5):
    """Single H2O: O at the grid centre, H at rOH and the given angle in the
    centre k-plane, fractional (unrounded) grid positions.  Returns
    (rows[n,5], els, screen)."""
    N2 = NN / 2
    ang = math.radians(angle_deg)
    rows = []
    els = []
    for s in (+1, -1):
        th = math.pi / 2 + s * ang / 2
        rows.append((N2 + rOH * math.cos(th) / h, N2 - rOH * math.sin(th) / h, N2, 1.0, 0.0))
        els.append("H")
        if s == +1:
            rows.append((N2, N2, N2, 2.0, 0.4))
            els.append("O")
    return np.array(rows, np.float64), els, NN * h


def _rand_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def water_geometry(rOH=1.809, angle_deg=104.5):
    """O at the origin, two H in the xy plane (bohr)."""
    ang = math.radians(angle_deg)
    return np.array([
        [0.0, 0.0, 0.0],
        [rOH * math.cos(ang / 2), rOH * math.sin(ang / 2), 0.0],
        [rOH * math.cos(ang / 2), -rOH * math.sin(ang / 2), 0.0],
    ])


def water_box(n_mol, box_bohr, NN, seed, min_OO_angstrom=2.6, margin_bohr=3.0, max_tries=200000):
    """n_mol waters at random positions/orientations in a cube of edge box_bohr.

    Rejection sampling of O positions inside [margin, box - margin]^3 with O-O >=
    ``min_OO_angstrom``; each molecule gets a uniformly random rotation.
    Returns (rows[n,5] in grid units, els, screen)."""
    rng = np.random.default_rng(seed)
    h = box_bohr / NN
    dmin = min_OO_angstrom * ANGSTROM
    O = np.empty((n_mol, 3))
    placed = 0
    tries = 0
    while placed < n_mol:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n_mol} waters with O-O >= {min_OO_angstrom} A")
        p = rng.uniform(margin_bohr, box_bohr - margin_bohr, size=3)
        if placed == 0 or ((O[:placed] - p) ** 2).sum(1).min() >= dmin * dmin:
            O[placed] = p
            placed += 1
    geo = water_geometry()
    rows, els = [], []
    for p in O:
        R = _rand_rotation(rng)
        for a, el in enumerate(("O", "H", "H")):
            x = p + R @ geo[a]
            Z, rc = ELEMENT[el]
            rows.append((x[0] / h, x[1] / h, x[2] / h, Z, rc))
            els.append(el)
    return np.array(rows, np.float64), els, box_bohr


def c1_h2o():
    """Plain method, single H2O on 96^3, h = 0.1 bohr."""
    rows, els, screen = h2o_single(96, 0.1)
    return {"NN": 96, "screen": screen, "rows": rows, "els": els, "mg_levels": 4}


def c2_waters():
    """10 H2O, seeded orientations, O-O >= 2.6 A, 12 A cube, 192^3, 5 multigrid levels."""
    rows, els, screen = water_box(10, 12.0 * ANGSTROM, 192, seed=1)
    return {"NN": 192, "screen": screen, "rows": rows, "els": els, "mg_levels": 5}


def c5_waters():
    """1000 H2O at 1 g/cm^3 in a 58.58 bohr cube (O-O >= 2.5 A, seed 4), 512^3
    (h = 0.1144 bohr), 7 multigrid levels: 3000 domains."""
    rows, els, screen = water_box(1000, 58.58, 512, seed=4, min_OO_angstrom=2.5, margin_bohr=2.0,
                                  max_tries=2000000)
    return {"NN": 512, "screen": screen, "rows": rows, "els": els, "mg_levels": 7}


# Tri-glycine NH2-Gly-Gly-Gly-COOH with one water (angstrom, element), the
# coordinate table of peptide_grow.html:22-56.
TRIGLYCINE = [
    ("N", 0.0, 0.0, 0.0), ("C", 1.47, 0.0, 0.0), ("C", 2.1, 1.23, 0.0), ("O", 1.6, 2.3, 0.0),
    ("N", 3.4, 1.2, 0.0), ("C", 4.1, 2.4, 0.0), ("C", 5.6, 2.3, 0.0), ("O", 6.2, 1.2, 0.0),
    ("N", 6.3, 3.4, 0.0), ("C", 7.7, 3.5, 0.0), ("C", 8.4, 4.7, 0.0), ("O", 7.9, 5.8, 0.0),
    ("O", 9.6, 4.6, 0.0), ("H", 10.2, 5.3, 0.0),
    ("H", -0.4, 0.8, 0.0), ("H", -0.4, -0.8, 0.0), ("H", 1.5, -0.5, 0.9), ("H", 1.5, -0.5, -0.9),
    ("H", 3.8, 0.3, 0.0), ("H", 3.8, 3.0, 0.9), ("H", 3.8, 3.0, -0.9), ("H", 5.8, 4.1, 0.0),
    ("H", 8.0, 2.6, 0.9), ("H", 8.0, 2.6, -0.9),
    ("O", 2.5, -1.5, 0.0), ("H", 1.8, -1.9, 0.0), ("H", 3.1, -2.0, 0.0),
]


def c4_peptide(NN=256, n_water=8, seed=3, jitter_angstrom=0.1, min_dist_angstrom=2.6):
    """Tri-glycine + its water (27 atoms) with a seeded +-0.1 A jitter, plus
    n_water waters placed at random orientations with every heavy atom >= 2.6 A
    from theirs; screen from peptide_grow.html:73-75 (extent + 15 A, clamped
    to [35, 90] bohr); NN = 256, 5 multigrid levels (coarsest 15^3)."""
    rng = np.random.default_rng(seed)
    els = [a[0] for a in TRIGLYCINE]
    X = np.array([a[1:] for a in TRIGLYCINE], np.float64)
    X += rng.uniform(-jitter_angstrom, jitter_angstrom, X.shape)
    X = (X - X.mean(0)) * ANGSTROM
    extent = (X.max(0) - X.min(0)).max()
    screen = float(min(90.0, max(35.0, extent + 15 * ANGSTROM)))
    centre = screen / 2
    heavy = [x for x, el in zip(X, els) if el != "H"]
    geo = water_geometry()
    dmin = min_dist_angstrom * ANGSTROM
    placed = 0
    tries = 0
    while placed < n_water:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place the waters")
        p = rng.uniform(-extent / 2 - 3.0, extent / 2 + 3.0, size=3)
        if np.abs(p).max() > centre - 3.0 or any(np.linalg.norm(p - q) < dmin for q in heavy):
            continue
        R = _rand_rotation(rng)
        for a, el in enumerate(("O", "H", "H")):
            X = np.vstack([X, p + R @ geo[a]])
            els.append(el)
        heavy.append(p)
        placed += 1
    h = screen / NN
    rows = []
    for el, x in zip(els, X):
        Z, rc = ELEMENT[el]
        rows.append(((centre + x[0]) / h, (centre + x[1]) / h, (centre + x[2]) / h, Z, rc))
    return {"NN": NN, "screen": screen, "rows": np.array(rows), "els": els, "mg_levels": 5}


# Molecule templates (angstrom) of the reference's grow pages, solvent waters
# dropped: formamide_grow.html:23-29, ethanol_grow.html:23-34,
# glycine_grow.html:21-33.  Each entry: element, x, y, z.  Torsion groups are
# (axis atom a, axis atom b, atoms rotated about the a-b bond).
TEMPLATES = {
    "h2o": {"atoms": [("O", 0.0, 0.0, 0.0), ("H", 0.9572 * math.cos(math.radians(52.25)),
                                             0.9572 * math.sin(math.radians(52.25)), 0.0),
                      ("H", 0.9572 * math.cos(math.radians(52.25)), -0.9572 * math.sin(math.radians(52.25)), 0.0)],
            "torsions": []},
    "formamide": {"atoms": [("C", 0.0, 0.0, 0.0), ("O", -0.6, 1.2, 0.0), ("N", 1.3, -0.3, 0.0),
                            ("H", -0.6, -0.9, 0.0), ("H", 1.8, -1.2, 0.0), ("H", 1.9, 0.5, 0.0)],
                  "torsions": [(0, 2, [4, 5])]},
    "ethanol": {"atoms": [("C", 0.0, 0.0, 0.0), ("C", 1.52, 0.0, 0.0), ("O", 2.1, 1.2, 0.5),
                          ("H", -0.4, 0.9, 0.4), ("H", -0.4, -0.5, 0.9), ("H", -0.4, -0.4, -0.9),
                          ("H", 1.9, -0.5, 0.9), ("H", 1.9, -0.5, -0.9), ("H", 3.0, 1.3, 0.5)],
                "torsions": [(0, 1, [3, 4, 5]), (1, 2, [8])]},
    "glycine": {"atoms": [("N", 0.0, 0.0, 0.0), ("C", 1.47, 0.0, 0.0), ("C", 2.0, 1.4, 0.0), ("O", 1.3, 2.4, 0.0),
                          ("O", 3.2, 1.5, 0.0), ("H", -0.4, 0.8, 0.3), ("H", -0.4, -0.8, -0.3),
                          ("H", 1.8, -0.5, 0.9), ("H", 1.8, -0.5, -0.9), ("H", 3.6, 2.3, 0.0)],
                "torsions": [(1, 2, [3, 4, 9]), (0, 1, [5, 6])]},
}


def _rotate_about(points, a, b, angle):
    """Rotate points about the axis through a and b by angle (Rodrigues)."""
    k = (b - a) / np.linalg.norm(b - a)
    p = points - a
    c, s = math.cos(angle), math.sin(angle)
    return a + p * c + np.cross(k, p) * s + np.outer(p @ k, k) * (1 - c)


def conformer(name, rng, max_torsion_deg=60.0):
    """Random conformer: template, every torsion group rotated by U(-60, 60)
    degrees, then a uniformly random overall rotation.  Bohr, centred."""
    t = TEMPLATES[name]
    els = [a[0] for a in t["atoms"]]
    X = np.array([[a[1], a[2], a[3]] for a in t["atoms"]]) * ANGSTROM
    for ia, ib, group in t["torsions"]:
        ang = math.radians(rng.uniform(-max_torsion_deg, max_torsion_deg))
        X[group] = _rotate_about(X[group], X[ia], X[ib], ang)
    X = X @ _rand_rotation(rng).T
    X -= X.mean(0)
    return els, X


def c3_conformers(n=64, seed=2, h=0.1, sizes=(64, 80, 96, 112, 128, 144, 160), margin_bohr=2.5):
    """Batched geometry scan: n seeded conformers of H2O, formamide, ethanol and
    glycine; each gets a grid NN drawn from ``sizes`` among those whose box
    (NN h) holds the molecule with ``margin_bohr`` on every side.  Returns a
    list of (rows[n,5], NN, screen, name)."""
    rng = np.random.default_rng(seed)
    names = ["h2o", "formamide", "ethanol", "glycine"]
    out = []
    for c in range(n):
        name = names[c % 4]
        els, X = conformer(name, rng)
        extent = (X.max(0) - X.min(0)).max()
        fits = [s for s in sizes if s * h >= extent + 2 * margin_bohr]
        NN = int(rng.choice(fits))
        screen = NN * h
        centre = screen / 2
        rows = []
        for el, x in zip(els, X):
            Z, rc = ELEMENT[el]
            rows.append(((centre + x[0]) / h, (centre + x[1]) / h, (centre + x[2]) / h, Z, rc))
        out.append((np.array(rows), NN, screen, name))
    return out


def gaussian_init(rows, NN, screen, label, seed=0, sigma_H=0.9, sigma_O=0.6, noise=1e-3):
    """Seeded Gaussian initial amplitudes per domain: exp(-r^2 / 2 sigma^2) about the
    domain's nucleus plus noise * N(0,1); zero on the faces.  ``label`` is the flat
    S^3 label map.  Returns a flat float32 S^3 array."""
    rng = np.random.default_rng(seed)
    S = NN + 1
    h = screen / NN
    x = np.arange(S) * h
    lab = np.asarray(label).reshape(S, S, S)
    U = np.zeros((S, S, S), np.float64)
    for n, r in enumerate(rows):
        sig = sigma_H if r[3] <= 1 else sigma_O
        m = lab == n
        if not m.any():
            continue
        ii, jj, kk = np.nonzero(m)
        d2 = (x[ii] - r[0] * h) ** 2 + (x[jj] - r[1] * h) ** 2 + (x[kk] - r[2] * h) ** 2
        U[ii, jj, kk] = np.exp(-d2 / (2 * sig * sig))
    U += noise * rng.standard_normal(U.shape)
    U[0, :, :] = U[-1, :, :] = 0
    U[:, 0, :] = U[:, -1, :] = 0
    U[:, :, 0] = U[:, :, -1] = 0
    return U.astype(np.float32).ravel()


def mf_h2o(N=200, bond_au=None, screen=12.0):
    """The mol_fast.js water scenario (mol_fast_h2o.html:18-46): O (Z = 3,
    rc = 0.5, triply occupied) and two H at 104.5 degrees, all fixed, on an
    N^3 grid of ``screen`` au; returns the page's USER_* configuration."""
    A = 1.8897
    b_oh = bond_au if bond_au is not None else 0.96 * A
    h = screen / N
    N2 = round(N / 2)
    ang = 104.5 * math.pi / 180
    hx, hy = b_oh * math.sin(ang / 2), b_oh * math.cos(ang / 2)
    nuclei = [dict(i=N2, j=N2, k=N2, Z=3, rc=0.5, tilt_y=0.0, fixed=True),
              dict(i=N2 + round(hx / h), j=N2 + round(hy / h), k=N2, Z=1, rc=0.0, fixed=True),
              dict(i=N2 - round(hx / h), j=N2 + round(hy / h), k=N2, Z=1, rc=0.0, fixed=True)]
    return {"USER_NN": N, "USER_SCREEN": screen, "USER_STEPS": 100000, "USER_NUCLEI": nuclei,
            "USER_NORM_TARGETS": [3, 1, 1]}


def hairpin_solvated():
    """The bent solvated beta-hairpin of the reference's recorded A100 run
    (recipe of cloud/run_hairpin.py:21-182; the page hairpin_bent_solvated.html):
    12 residues on a U-shaped path in the k = N2 plane, backbone N, CA, C with
    their H and the carbonyl O, a C-terminal OH, and a shell of waters on a
    2.8 A lattice within 8 A of the protein (none closer than 2.5 A); 300^3
    cells over an 80-bohr box, positions rounded to whole grid nodes with
    Python's round().  Returns the RealQMSolver config dict of that run (10000
    steps, reports every 1000, dynamics on with forceScale 500, fold atoms
    N-terminus / middle CA / C-terminus)."""
    NN, screen = 300, 80.0
    h = screen / NN
    N2 = NN // 2
    A = ANGSTROM
    atoms, xy = [], []

    def put(x, y, el):
        Z, rc = ELEMENT[el]
        atoms.append({"i": round(N2 + x / h), "j": round(N2 + y / h), "k": N2, "Z": Z, "el": el, "rc": rc})
        xy.append((x, y))

    n_res, rise, zig = 12, 3.4 * A, 0.5 * A
    # strand direction per residue: the two strands 30 degrees apart, blended
    # through the turn by tanh((r - 5.5) / 2.5)
    half_open = math.radians(30.0) / 2
    a_start = -(math.pi - (math.pi / 2 + half_open))
    a_end = math.pi / 2 - half_open
    ang = []
    for r in range(n_res):
        w = 0.5 * (1.0 + math.tanh((r - 5.5) / 2.5))
        ang.append(a_start * (1.0 - w) + a_end * w)
    path = [(0.0, 0.0)]
    for r in range(1, n_res):
        mid = 0.5 * (ang[r - 1] + ang[r])
        path.append((path[-1][0] + rise * math.cos(mid), path[-1][1] + rise * math.sin(mid)))
    cx0 = sum(p[0] for p in path) / n_res
    cy0 = sum(p[1] for p in path) / n_res
    backbone = []
    for r in range(n_res):
        px, py = path[r][0] - cx0, path[r][1] - cy0
        tx, ty = math.cos(ang[r]), math.sin(ang[r])
        nx, ny = -ty, tx
        sgn = 1 if r % 2 == 0 else -1
        lift = 15.0  # the hairpin sits 15 bohr up in j
        n_at = (px - 0.35 * rise * tx + sgn * zig * 0.6 * nx, py - 0.35 * rise * ty + sgn * zig * 0.6 * ny + lift)
        ca_at = (px, py - sgn * zig * 0.4 * ny + lift)
        c_at = (px + 0.35 * rise * tx + sgn * zig * 0.3 * nx, py + 0.35 * rise * ty + sgn * zig * 0.3 * ny + lift)
        backbone.append((n_at, ca_at, c_at))

    def unit_normal(p, q):
        dx, dy = q[0] - p[0], q[1] - p[1]
        L = math.hypot(dx, dy)
        return -dy / L, dx / L

    d_nh, d_ch, d_co = 1.01 * A, 1.09 * A, 1.24 * A
    for r, (n_at, ca_at, c_at) in enumerate(backbone):
        pnx, pny = unit_normal(n_at, ca_at)
        pcx, pcy = unit_normal(ca_at, c_at)
        put(n_at[0], n_at[1], "N")
        s_nh = -1 if r < 6 else 1
        put(n_at[0] + s_nh * pnx * d_nh, n_at[1] + s_nh * pny * d_nh, "H")
        if r == 0:  # the N-terminal NH2
            put(n_at[0] - s_nh * pnx * d_nh, n_at[1] - s_nh * pny * d_nh, "H")
        put(ca_at[0], ca_at[1], "C")
        put(ca_at[0] + pcx * d_ch, ca_at[1] + pcy * d_ch, "H")
        put(ca_at[0] - pcx * d_ch, ca_at[1] - pcy * d_ch, "H")
        put(c_at[0], c_at[1], "C")
        even = r % 2 == 0
        s_co = (1 if even else -1) if r < 6 else (-1 if even else 1)
        put(c_at[0] + s_co * pcx * d_co, c_at[1] + s_co * pcy * d_co, "O")
    # C-terminal OH along the last CA -> C bond
    ca_l, c_l = backbone[-1][1], backbone[-1][2]
    ux, uy = c_l[0] - ca_l[0], c_l[1] - ca_l[1]
    L = math.hypot(ux, uy)
    ux, uy = ux / L, uy / L
    ox, oy = c_l[0] + ux * 1.34 * A, c_l[1] + uy * 1.34 * A
    put(ox, oy, "O")
    put(ox + ux * 0.97 * A, oy + uy * 0.97 * A, "H")
    n_protein = len(atoms)

    # waters on a lattice: kept if no protein atom is within 2.5 A and one is
    # within 8 A; orientation from a fixed hash of the lattice point
    d_oh, half = 0.96 * A, math.radians(52.25)
    step, near, shell = 2.8 * A, 2.5 * A, 8.0 * A
    lim = screen / 2 - 3
    prot = xy[:n_protein]
    wx = -lim
    while wx <= lim:
        wy = -lim
        while wy <= lim:
            clash, close = False, False
            for (px, py) in prot:
                d2 = (wx - px) ** 2 + (wy - py) ** 2
                if d2 < near * near:
                    clash = True
                    break
                if d2 < shell * shell:
                    close = True
            if close and not clash:
                phi = (wx * 7.3 + wy * 11.1) % (2 * math.pi)
                put(wx, wy, "O")
                put(wx + d_oh * math.cos(phi + half), wy + d_oh * math.sin(phi + half), "H")
                put(wx + d_oh * math.cos(phi - half), wy + d_oh * math.sin(phi - half), "H")
            wy += step
        wx += step
    return {"gridN": NN, "screen": screen, "atoms": atoms, "steps": 10000, "report_interval": 1000,
            "dynamics": True, "forceScale": 500.0, "fold_atoms": [0, 45, 83], "nProtein": n_protein}
