"""Seeded synthetic receptors, ligands and poses in the reference data model.

The container has no RDKit/AmberTools, so every test and benchmark input is
generated here and is a valid reference input (data.receptor_json /
data.ligand_json write the reference JSON).  Sizes follow the documented
workloads: a 2,000-heavy-atom receptor in a 30 A-radius sphere with a 7 A
pocket and 8 rotatable hydroxyl hydrogens; ligands of 10-60 heavy atoms with
0-15 rotatable bonds and 0-3 rings; elements C/N/O/S/F at 70/12/14/2/2 %.

Force-field terms are built from the generated topology with the reference
preparation rules: non-bonded 1-4 / all-pair lists exclude 1-2 and 1-3 pairs
(pxdock/parser/ffdata.py:179-227) and receptor terms are pruned to those that
touch a movable atom (pxdock/parser/receptor.py:230-265).
"""
import math
from itertools import combinations

import numpy as np

from .data import FragTree, Leaves, Molecule

# GAFF2 type indices (order of the 99-entry vdW table)
T = {"c": 0, "c3": 4, "ca": 5, "hn": 28, "ho": 29, "f": 34, "n": 38, "n3": 41, "nb": 44,
     "nh": 49, "o": 72, "oh": 73, "os": 76, "sh": 93, "ss": 94}
# receptor LJ by element (AMBER-like)
REC_LJ = {"C": (3.40, 0.086), "N": (3.25, 0.17), "O": (2.96, 0.21), "S": (3.56, 0.25),
          "HN": (1.07, 0.0157), "HO": (0.0, 0.0)}
ATOMIC = {"C": 6, "N": 7, "O": 8, "S": 16, "F": 9, "H": 1}


def _unit(v):
    return v / np.linalg.norm(v)


def _rand_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _perp(d, rng):
    v = _rand_unit(rng)
    v = v - d * (v @ d)
    return v / np.linalg.norm(v)


def quat_to_rot(q):
    """get_rotation_matrix(quaternion) (engine/src/lib/math.cpp:22-42)."""
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (a * c + b * d)],
                     [2 * (a * d + b * c), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (a * b + c * d), a * a - b * b - c * c + d * d]])


def random_quaternion(rng, n=None):
    """Uniform unit quaternion from three uniforms (random.cpp:29-40)."""
    u = rng.random((3,) if n is None else (n, 3))
    u1, u2, u3 = u[..., 0], u[..., 1] * 2 * math.pi, u[..., 2] * 2 * math.pi
    s1, s2 = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([s1 * np.sin(u2), s1 * np.cos(u2), s2 * np.sin(u3), s2 * np.cos(u3)], -1)


def axis_rot(k, t):
    """Rodrigues matrix (engine/src/lib/math.cpp:44-60)."""
    x, y, z = k
    s, c = math.sin(t), math.cos(t)
    o = 1 - c
    return np.array([[c + x * x * o, x * y * o - z * s, y * s + x * z * o],
                     [z * s + x * y * o, c + y * y * o, y * z * o - x * s],
                     [x * z * o - y * s, x * s + y * z * o, c + z * z * o]])


# --------------------------------------------------------------------------
# topology -> force field


def _neighbors(n, bonds):
    nb = [[] for _ in range(n)]
    for i, j in bonds:
        nb[i].append(j)
        nb[j].append(i)
    return nb


def _angle(x, i, j, k):
    v0, v1 = x[i] - x[j], x[k] - x[j]
    return math.degrees(math.atan2(np.linalg.norm(np.cross(v0, v1)), v0 @ v1))


def _ff_from_topology(x, bonds, ring_atoms, rng):
    """Bonds, angles, propers, impropers and the non-bonded lists."""
    n = len(x)
    nb = _neighbors(n, bonds)
    bond_ij, bond_k, bond_r0 = [], [], []
    for i, j in bonds:
        bond_ij.append((i, j))
        bond_k.append(rng.uniform(300.0, 700.0))
        bond_r0.append(np.linalg.norm(x[i] - x[j]) * rng.uniform(0.98, 1.02))
    angle_ijk, angle_k, angle_t0 = [], [], []
    for j in range(n):
        for i, k in combinations(sorted(nb[j]), 2):
            angle_ijk.append((i, j, k))
            angle_k.append(rng.uniform(50.0, 100.0))
            angle_t0.append(math.radians(_angle(x, i, j, k) * rng.uniform(0.97, 1.03)))
    propers = []
    for j, k in bonds:
        for i in nb[j]:
            if i == k:
                continue
            for l in nb[k]:
                if l == j or l == i:
                    continue
                propers.append((i, j, k, l))
    dih_ijkl, dih_n, dih_phase, dih_k = [], [], [], []
    impropers = []
    for c in range(n):
        if c in ring_atoms and len(nb[c]) == 3:
            a, b, d = nb[c]
            impropers.append((c, a, b, d))
    for q in impropers:
        dih_ijkl.append(q)
        dih_n.append([2, 0, 0, 0, 0, 0])
        dih_phase.append([math.pi, 0, 0, 0, 0, 0])
        dih_k.append([1.1, 0, 0, 0, 0, 0])
    for q in propers:
        nt = int(rng.integers(1, 4))
        pers = sorted(rng.choice([1, 2, 3], size=nt, replace=False).tolist())
        dih_ijkl.append(q)
        dih_n.append(pers + [0] * (6 - nt))
        dih_phase.append([math.pi * float(rng.integers(0, 2)) for _ in range(nt)] + [0] * (6 - nt))
        dih_k.append([rng.uniform(-0.3, 0.6) for _ in range(nt)] + [0] * (6 - nt))
    p12 = {tuple(sorted(b)) for b in bonds}
    p13 = {tuple(sorted((a[0], a[2]))) for a in angle_ijk}
    p14 = {tuple(sorted((q[0], q[3]))) for q in propers} - p12 - p13
    allp = [(i, j) for i in range(n) for j in range(i + 1, n)]
    excl = p12 | p13 | p14
    pairs = [p for p in allp if p not in excl]
    i2 = lambda v: np.asarray(v, np.int32).reshape(-1, 2)
    return dict(
        bond_ij=i2(bond_ij), bond_k=np.asarray(bond_k), bond_r0=np.asarray(bond_r0),
        angle_ijk=np.asarray(angle_ijk, np.int32).reshape(-1, 3), angle_k=np.asarray(angle_k),
        angle_t0=np.asarray(angle_t0), n_impropers=len(impropers),
        dih_ijkl=np.asarray(dih_ijkl, np.int32).reshape(-1, 4),
        dih_n=np.asarray(dih_n, np.int32).reshape(-1, 6),
        dih_phase=np.asarray(dih_phase, float).reshape(-1, 6),
        dih_k=np.asarray(dih_k, float).reshape(-1, 6),
        pairs14=i2(sorted(p14)), pairs=i2(pairs))


# --------------------------------------------------------------------------
# ligands


def _grow_ligand(rng, n_heavy, n_rings):
    """Random tree of rings and chain atoms embedded in 3D by growth."""
    elems, ring_of, x = [], [], []
    bonds = []
    rings = []
    cap = []

    def clash(p, skip):
        for a, q in enumerate(x):
            if a in skip:
                continue
            if np.linalg.norm(p - q) < 2.45:
                return True
        return False

    def direction(parent, rng):
        nbr = [x[b] - x[parent] for (a, b) in bonds if a == parent] + \
              [x[a] - x[parent] for (a, b) in bonds if b == parent]
        for _ in range(60):
            d = _rand_unit(rng)
            if all(_unit(v) @ d < -0.2 for v in nbr):
                return d
        return None

    ring_sizes = [6 if rng.random() < 0.75 else 5 for _ in range(n_rings)]
    while sum(ring_sizes) > n_heavy - 1 and ring_sizes:
        ring_sizes.pop()
    n_chain = n_heavy - sum(ring_sizes)
    chain_elems = list(rng.choice(["C", "N", "O", "S", "F"], size=n_chain,
                                  p=[0.70, 0.12, 0.14, 0.02, 0.02]))
    units = [("ring", s) for s in ring_sizes] + [("atom", e) for e in chain_elems]
    # start from a carbon-ish unit so growth has capacity
    rng.shuffle(units)
    units.sort(key=lambda u: u[0] == "atom" and u[1] in ("F", "O", "S"))
    capacity = {"C": 4, "N": 3, "O": 2, "S": 2, "F": 1}

    def add_ring(size, attach_pos, d):
        r = 1.39 / (2 * math.sin(math.pi / size))
        center = attach_pos + r * d
        v = _perp(d, rng)
        w = np.cross(d, v)
        idx = []
        pts = []
        for k in range(size):
            th = math.pi + 2 * math.pi * k / size
            pts.append(center + r * (math.cos(th) * d + math.sin(th) * v) * 1.0 + 0 * w)
        return pts

    for u_i, (kind, val) in enumerate(units):
        placed = False
        for _attempt in range(40):
            if not x:
                if kind == "ring":
                    pts = add_ring(val, np.zeros(3), _rand_unit(rng))
                else:
                    pts = [np.zeros(3)]
                parent = None
            else:
                cands = [a for a in range(len(x)) if cap[a] > 0]
                if not cands:
                    break
                parent = int(rng.choice(cands))
                d = direction(parent, rng)
                if d is None:
                    continue
                bl = rng.uniform(1.33, 1.54)
                start = x[parent] + bl * d
                pts = add_ring(val, start, d) if kind == "ring" else [start]
            skip = set([parent]) if parent is not None else set()
            if parent is not None:
                skip |= {b for (a, b) in bonds if a == parent} | {a for (a, b) in bonds if b == parent}
            if any(clash(p, skip) for p in pts):
                continue
            if kind == "ring" and len(pts) > 1:
                inner = [np.linalg.norm(pts[a] - pts[b]) for a in range(len(pts))
                         for b in range(a + 1, len(pts))]
                if min(inner) < 1.2:
                    continue
            base = len(x)
            if kind == "ring":
                rid = len(rings)
                members = list(range(base, base + val))
                rings.append(members)
                for k, p in enumerate(pts):
                    el = "N" if (k > 0 and rng.random() < 0.12) else "C"
                    elems.append(el)
                    ring_of.append(rid)
                    x.append(p)
                    cap.append(1 if el == "C" else 0)
                for k in range(val):
                    bonds.append((base + k, base + (k + 1) % val))
            else:
                elems.append(val)
                ring_of.append(-1)
                x.append(pts[0])
                cap.append(capacity[val])
            if parent is not None:
                bonds.append((parent, base))
                cap[parent] -= 1
                cap[base] -= 1
            placed = True
            break
        if not placed:
            return None
    return elems, ring_of, np.array(x), bonds, rings


def _fragments(n, bonds, rot):
    """Connected components after cutting the rotatable bonds."""
    cut = {tuple(sorted(b)) for b in rot}
    nb = [[] for _ in range(n)]
    for i, j in bonds:
        if tuple(sorted((i, j))) in cut:
            continue
        nb[i].append(j)
        nb[j].append(i)
    comp = [-1] * n
    frags = []
    for s in range(n):
        if comp[s] >= 0:
            continue
        stack, members = [s], []
        comp[s] = len(frags)
        while stack:
            a = stack.pop()
            members.append(a)
            for b in nb[a]:
                if comp[b] < 0:
                    comp[b] = len(frags)
                    stack.append(b)
        frags.append(sorted(members))
    return frags, comp


def _tree_from_fragments(frags, comp, rot):
    """Fragment tree rooted at the fragment-graph centre (as geometry/tree.py:39-48),
    fragment ids in BFS order, children sorted by id."""
    nf = len(frags)
    adj = {f: [] for f in range(nf)}
    edge = {}
    for i, j in rot:
        a, b = comp[i], comp[j]
        adj[a].append(b)
        adj[b].append(a)
        edge[(a, b)] = (i, j)
        edge[(b, a)] = (j, i)

    def bfs(src):
        order, par = [src], {src: -1}
        k = 0
        while k < len(order):
            for c in sorted(adj[order[k]]):
                if c not in par:
                    par[c] = order[k]
                    order.append(c)
            k += 1
        return order, par

    far1 = bfs(0)[0][-1]
    order, par = bfs(far1)
    far2 = order[-1]
    path = [far2]
    while par[path[-1]] != -1:
        path.append(par[path[-1]])
    root = path[len(path) // 2]
    order, par = bfs(root)
    newid = {f: i for i, f in enumerate(order)}
    new_frags = [frags[f] for f in order]
    parent = np.array([newid[par[f]] if par[f] >= 0 else -1 for f in order], np.int32)
    perm = np.array(sum(new_frags, []), np.int32)
    rev = np.empty(len(perm), np.int32)
    rev[perm] = np.arange(len(perm))
    sizes = np.array([len(f) for f in new_frags])
    end = np.cumsum(sizes).astype(np.int32)
    start = (end - sizes).astype(np.int32)
    bond = np.zeros((nf, 2), np.int32)
    children = {i: [] for i in range(nf)}
    for f in range(nf):
        p = int(parent[f])
        if p < 0:
            continue
        a, b = edge[(order[p], order[f])]
        bond[f] = (rev[a], rev[b])
        children[p].append(f)
    pre = []

    def visit(f):
        for c in sorted(children[f]):
            pre.append(c)
            visit(c)

    visit(0)
    return FragTree(perm=perm, frag_start=start, frag_end=end, frag_parent=parent,
                    frag_bond=bond, tors_frag=np.array(pre, np.int32))


def make_ligand(seed, n_heavy=30, n_torsions=6, n_rings=1, n_poses=1):
    """A ligand with its reference conformation as pose 0."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        g = _grow_ligand(rng, n_heavy, n_rings)
        if g is not None:
            break
    else:
        raise RuntimeError("ligand growth failed")
    elems, ring_of, xh, hbonds, rings = g
    nh = len(elems)
    nbh = _neighbors(nh, hbonds)
    # polar hydrogens on N/O with spare valence
    elems = list(elems)
    x = [p for p in xh]
    bonds = list(hbonds)
    hpair = []
    for a in range(nh):
        if elems[a] in ("N", "O") and ring_of[a] < 0 and len(nbh[a]) <= (2 if elems[a] == "N" else 1):
            dirs = [_unit(x[b] - x[a]) for b in nbh[a]]
            for _ in range(40):
                d = _rand_unit(rng)
                if all(v @ d < -0.25 for v in dirs):
                    break
            h = len(x)
            x.append(x[a] + 1.01 * d)
            elems.append("H")
            ring_of.append(-1)
            bonds.append((a, h))
            hpair.append((a, h))
    x = np.array(x)
    n = len(x)
    nb = _neighbors(n, bonds)
    ring_set = {a for r in rings for a in r}
    # GAFF2-style types
    vtype = np.zeros(n, np.int32)
    for a, e in enumerate(elems):
        heavy_nb = [b for b in nb[a] if elems[b] != "H"]
        if e == "C":
            vtype[a] = T["ca"] if a in ring_set else (T["c"] if any(elems[b] == "O" and len(nb[b]) == 1 for b in nb[a]) else T["c3"])
        elif e == "N":
            vtype[a] = T["nb"] if a in ring_set else (T["nh"] if any(b in ring_set for b in heavy_nb) else T["n3"])
        elif e == "O":
            vtype[a] = T["oh"] if any(elems[b] == "H" for b in nb[a]) else (T["o"] if len(heavy_nb) == 1 else T["os"])
        elif e == "S":
            vtype[a] = T["sh"] if len(heavy_nb) == 1 else T["ss"]
        elif e == "F":
            vtype[a] = T["f"]
        else:
            vtype[a] = T["hn"] if elems[nb[a][0]] == "N" else T["ho"]
    # rotatable bonds: chain single bonds between non-terminal heavy atoms
    cand = []
    for i, j in hbonds:
        if ring_of[i] >= 0 and ring_of[i] == ring_of[j]:
            continue
        if len([b for b in nbh[i]]) < 2 or len([b for b in nbh[j]]) < 2:
            continue
        cand.append((i, j))
    rng.shuffle(cand)
    rot = cand[:n_torsions]
    frags, comp = _fragments(n, bonds, rot)
    tree = _tree_from_fragments(frags, comp, rot)
    ff = _ff_from_topology(x, bonds, ring_set, rng)
    q = rng.normal(0.0, 0.3, size=n)
    q -= q.mean()  # zero net charge (pxdock/parser/ligand_ffdata.py:284-296)
    # pscore sets
    hydro = [a for a in range(n) if (elems[a] == "C" and not any(elems[b] in ("N", "O") for b in nb[a])) or elems[a] == "S"]
    cation = [a for a in range(n) if elems[a] == "N" and vtype[a] == T["n3"] and len([b for b in nb[a] if elems[b] != "H"]) == 1 and rng.random() < 0.5]
    anion = [a for a in range(n) if elems[a] == "O" and vtype[a] == T["o"] and rng.random() < 0.4]
    charged = set(cation) | set(anion)
    don_c = [(a, h) for a, h in hpair if a in charged]
    don_n = [(a, h) for a, h in hpair if a not in charged]
    acc = [a for a in range(n) if elems[a] == "O" or (elems[a] == "N" and a in ring_set)]
    acc_c = [a for a in acc if a in charged]
    acc_n = [a for a in acc if a not in charged]
    ring5 = [r for r in rings if len(r) == 5]
    ring6 = [r for r in rings if len(r) == 6]
    # torsion-strain penalties on about half of the rotatable bonds
    ts_coef, ts_off, ts_ijkl = [], [0], []
    for i, j in rot:
        if rng.random() < 0.5:
            continue
        a_opts = [a for a in nb[i] if a != j]
        d_opts = [d for d in nb[j] if d != i]
        if not a_opts or not d_opts:
            continue
        nm = int(rng.integers(1, 3))
        matches = [(int(rng.choice(a_opts)), i, j, int(rng.choice(d_opts))) for _ in range(nm)]
        ts_coef.append(rng.uniform(0.0, 0.5, size=4))
        ts_ijkl.extend(matches)
        ts_off.append(len(ts_ijkl))
    ligand = Molecule(
        xyz=x[None].copy(), charge=q, vdw_type=vtype,
        hydrophobic=np.array(hydro, np.int32), cation=np.array(cation, np.int32),
        anion=np.array(anion, np.int32),
        ring5=np.array(ring5, np.int32).reshape(-1, 5), ring6=np.array(ring6, np.int32).reshape(-1, 6),
        don_charged=np.array(don_c, np.int32).reshape(-1, 2),
        don_neutral=np.array(don_n, np.int32).reshape(-1, 2),
        acc_charged=np.array(acc_c, np.int32), acc_neutral=np.array(acc_n, np.int32),
        tstrain_coef=np.array(ts_coef, float).reshape(-1, 4), tstrain_off=np.array(ts_off, np.int32),
        tstrain_ijkl=np.array(ts_ijkl, np.int32).reshape(-1, 4),
        rotatable_bonds=np.array(rot, np.int32).reshape(-1, 2),
        all_bonds=np.array(bonds, np.int32).reshape(-1, 2),
        atomic_numbers=np.array([ATOMIC[e] for e in elems], np.int32),
        molecular_weight=float(sum({"C": 12.011, "N": 14.007, "O": 15.999, "S": 32.06, "F": 18.998, "H": 1.008}[e] for e in elems)),
        smiles=f"synthetic-{seed}", tree=tree, **ff)
    return ligand


# --------------------------------------------------------------------------
# receptors


def make_receptor(seed=1, n_heavy=2000, radius=30.0, pocket_offset=15.0, pocket_radius=7.0,
                  min_sep=2.8, n_leaves=8, n_ring6=20, n_ring5=6):
    """Random-sequential-addition receptor with a carved pocket (see module doc)."""
    rng = np.random.default_rng(seed)
    pocket = np.array([pocket_offset, 0.0, 0.0])
    cell = min_sep
    grid = {}
    pts, elems, ring_members = [], [], []

    def key(p):
        return tuple((p // cell).astype(int))

    def ok(p, own=()):
        if np.linalg.norm(p) > radius or np.linalg.norm(p - pocket) < pocket_radius:
            return False
        k = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for a in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if a in own:
                            continue
                        if np.linalg.norm(pts[a] - p) < min_sep:
                            return False
        return True

    def add(p, e):
        grid.setdefault(key(p), []).append(len(pts))
        pts.append(p)
        elems.append(e)

    def rand_in_sphere():
        while True:
            p = rng.uniform(-radius, radius, 3)
            if np.linalg.norm(p) <= radius:
                return p

    for size, count in ((6, n_ring6), (5, n_ring5)):
        placed = 0
        while placed < count:
            c = rand_in_sphere()
            d = _rand_unit(rng)
            v = _perp(d, rng)
            w = np.cross(d, v)
            r = 1.39 / (2 * math.sin(math.pi / size))
            ring = [c + r * (math.cos(2 * math.pi * k / size) * v + math.sin(2 * math.pi * k / size) * w)
                    for k in range(size)]
            if all(ok(p) for p in ring):
                base = len(pts)
                for k, p in enumerate(ring):
                    add(p, "N" if (size == 5 and k == 0) else "C")
                ring_members.append(list(range(base, base + size)))
                placed += 1
    ring_atoms = {a for r in ring_members for a in r}
    while len(pts) < n_heavy:
        p = rand_in_sphere()
        if ok(p):
            u = rng.random()
            e = "C" if u < 0.63 else ("N" if u < 0.80 else ("O" if u < 0.99 else "S"))
            add(p, e)
    heavy = np.array(pts)
    nh = len(heavy)
    # rotatable hydroxyls: O atoms in the pocket shell
    dist_p = np.linalg.norm(heavy - pocket, axis=1)
    o_shell = [a for a in range(nh) if elems[a] == "O" and dist_p[a] < pocket_radius + 4.0]
    rng.shuffle(o_shell)
    leaf_o = o_shell[:n_leaves]
    # polar H on 50 % of N/O (always on the leaf oxygens)
    xs = [p for p in heavy]
    el = list(elems)
    h_of = {}
    leaves_info = []
    for a in range(nh):
        if el[a] not in ("N", "O") or a in ring_atoms:
            continue
        if a not in leaf_o and rng.random() >= 0.5:
            continue
        h = len(xs)
        if a in leaf_o:
            d2 = np.linalg.norm(heavy - heavy[a], axis=1)
            d2[a] = np.inf
            X = int(np.argmin(d2))
            d2[X] = np.inf
            Y = int(np.argmin(d2))
            ax = _unit(heavy[a] - heavy[X])
            v = _perp(ax, rng)
            ang = math.radians(108.5)
            hp = heavy[a] + 0.96 * (math.cos(math.pi - ang) * ax + math.sin(math.pi - ang) * v)
            leaves_info.append((X, a, h, Y))
        else:
            hp = heavy[a] + 1.0 * _rand_unit(rng)
        xs.append(hp)
        el.append("HO" if el[a] == "O" else "HN")
        h_of[a] = h
    x = np.array(xs)
    n = len(x)
    sigma = np.array([REC_LJ[e][0] for e in el])
    epsilon = np.array([REC_LJ[e][1] for e in el])
    q = rng.normal(0.0, 0.25, size=n)
    q -= q.mean()
    movable = [h for (_, _, h, _) in leaves_info]
    mov_set = set(movable)
    bond_ij, bond_k, bond_r0 = [], [], []
    angle_ijk, angle_k, angle_t0 = [], [], []
    dih = []
    ex12, ex13, p14 = set(), set(), set()
    for X, O, H, Y in leaves_info:
        bond_ij.append((O, H))
        bond_k.append(553.0)
        bond_r0.append(0.96)
        angle_ijk.append((X, O, H))
        angle_k.append(47.0)
        angle_t0.append(math.radians(108.5))
        dih.append((Y, X, O, H))
        ex12.add(tuple(sorted((O, H))))
        ex13.add(tuple(sorted((X, H))))
        p14.add(tuple(sorted((Y, H))))
    pairs = set()
    for m in movable:
        for j in range(n):
            if j == m:
                continue
            pr = tuple(sorted((m, j)))
            if pr in ex12 or pr in ex13 or pr in p14:
                continue
            pairs.add(pr)
    # pscore sets
    N_idx = [a for a in range(nh) if el[a] == "N" and a not in ring_atoms]
    O_idx = [a for a in range(nh) if el[a] == "O"]
    cation = [a for a in N_idx if rng.random() < 0.1]
    anion = [a for a in O_idx if rng.random() < 0.1 and a not in leaf_o]
    charged = set(cation) | set(anion)
    don = [(a, h) for a, h in h_of.items()]
    acc = O_idx + [a for a in N_idx if a not in h_of]
    rec = Molecule(
        xyz=x, charge=q, sigma=sigma, epsilon=epsilon,
        bond_ij=np.array(bond_ij, np.int32).reshape(-1, 2), bond_k=np.array(bond_k),
        bond_r0=np.array(bond_r0),
        angle_ijk=np.array(angle_ijk, np.int32).reshape(-1, 3), angle_k=np.array(angle_k),
        angle_t0=np.array(angle_t0), n_impropers=0,
        dih_ijkl=np.array(dih, np.int32).reshape(-1, 4),
        dih_n=np.tile(np.array([3, 0, 0, 0, 0, 0], np.int32), (len(dih), 1)),
        dih_phase=np.zeros((len(dih), 6)),
        dih_k=np.tile(np.array([0.16, 0, 0, 0, 0, 0]), (len(dih), 1)),
        pairs14=np.array(sorted(p14), np.int32).reshape(-1, 2),
        pairs=np.array(sorted(pairs), np.int32).reshape(-1, 2),
        hydrophobic=np.array([a for a in range(nh) if el[a] in ("C", "S")], np.int32),
        cation=np.array(cation, np.int32), anion=np.array(anion, np.int32),
        ring5=np.array([r for r in ring_members if len(r) == 5], np.int32).reshape(-1, 5),
        ring6=np.array([r for r in ring_members if len(r) == 6], np.int32).reshape(-1, 6),
        don_charged=np.array([d for d in don if d[0] in charged], np.int32).reshape(-1, 2),
        don_neutral=np.array([d for d in don if d[0] not in charged], np.int32).reshape(-1, 2),
        acc_charged=np.array([a for a in acc if a in charged], np.int32),
        acc_neutral=np.array([a for a in acc if a not in charged], np.int32),
        atomic_numbers=np.array([ATOMIC[e[0]] for e in el], np.int32))
    origin, axis, off, atoms, axis_atoms, frs = [], [], [0], [], [], []
    for X, O, H, Y in leaves_info:
        origin.append(x[X])
        axis.append(_unit(x[O] - x[X]))
        atoms.append(H)
        off.append(len(atoms))
        axis_atoms.append((X, O))
        frs.append([O, H])
    L = len(leaves_info)
    rec.leaves = Leaves(origin=np.array(origin).reshape(L, 3), axis=np.array(axis).reshape(L, 3),
                        atom_off=np.array(off, np.int32), atoms=np.array(atoms, np.int32),
                        axis_atoms=np.array(axis_atoms, np.int32).reshape(L, 2), fragments=frs,
                        bond_order=np.arange(L, dtype=np.int32), keys=[str(k) for k in range(L)])
    rec.pocket_center = pocket
    return rec


# --------------------------------------------------------------------------
# poses


def ligand_apply(tree, xyz, center, rot, tors):
    """free_ligand::apply_parameters (engine/src/core/data.cpp:522-641) in numpy,
    same per-fragment restatement as the CPU checker (see its comment)."""
    n = len(xyz)
    cp = xyz[tree.perm]
    nf = tree.n_frags
    theta = np.zeros(nf)
    theta[tree.tors_frag] = tors
    axes = np.zeros((nf, 3))
    for f in range(nf):
        if tree.frag_parent[f] >= 0:
            a = cp[tree.frag_bond[f, 1]] - cp[tree.frag_bond[f, 0]]
            axes[f] = a / np.linalg.norm(a)
    c = cp.mean(axis=0)
    loc = np.zeros_like(cp)
    for f in range(nf):
        s, e = tree.frag_start[f], tree.frag_end[f]
        par = tree.frag_parent[f]
        if par < 0:
            loc[s:e] = cp[s:e] - c
            continue
        mine = tree.frag_bond[f, 1]
        loc[s:e] = cp[s:e] - cp[mine]
        ref = c if tree.frag_parent[par] < 0 else cp[tree.frag_bond[par, 1]]
        loc[mine] = cp[mine] - ref
    out = np.zeros_like(cp)
    M = [None] * nf
    O = [None] * nf
    for f in range(nf):
        if tree.frag_parent[f] < 0:
            M[f], O[f] = rot, np.asarray(center, float)
            s, e = tree.frag_start[f], tree.frag_end[f]
            out[s:e] = loc[s:e] @ rot.T + center
    for f in tree.tors_frag:
        par = tree.frag_parent[f]
        mine = tree.frag_bond[f, 1]
        out[mine] = M[par] @ loc[mine] + O[par]
        M[f] = axis_rot(M[par] @ axes[f], theta[f]) @ M[par]
        O[f] = out[mine]
        s, e = tree.frag_start[f], tree.frag_end[f]
        keep = out[mine].copy()
        out[s:e] = loc[s:e] @ M[f].T + O[f]
        out[mine] = keep
    res = np.empty_like(out)
    res[tree.perm] = out
    return res


def random_poses(ligand, box, n_poses, seed):
    """Poses with centre uniform in the box, a uniform unit quaternion and
    torsions uniform in [-pi, pi) applied to the reference conformation."""
    rng = np.random.default_rng(seed)
    lo = np.array(box.init_xyz)
    hi = np.array(box.oppo_xyz)
    out = np.zeros((n_poses, ligand.n_atoms, 3))
    x0 = ligand.xyz[0]
    for p in range(n_poses):
        c = rng.uniform(lo, hi)
        R = quat_to_rot(random_quaternion(rng))
        t = rng.uniform(-math.pi, math.pi, size=ligand.tree.n_torsions)
        out[p] = ligand_apply(ligand.tree, x0, c, R, t)
    return out


def ligand_set(n, seed0, heavy=(10, 60), tors=(0, 15), rings=(0, 3)):
    """n ligands with seeds seed0+i and sizes uniform in the given ranges."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i + 7919)
        nh = int(rng.integers(heavy[0], heavy[1] + 1))
        nt = int(rng.integers(tors[0], tors[1] + 1))
        nr = int(rng.integers(rings[0], rings[1] + 1))
        out.append(make_ligand(seed0 + i, n_heavy=nh, n_torsions=nt, n_rings=nr))
    return out
