"""Seeded synthetic decks for the linear-static workloads.

Node and element numbering follow the reference's deck generators:
CHEXA8 blocks use id = i*(ny+1)*(nz+1) + j*(nz+1) + k + 1 with corners
(-,-,-),(+,-,-),(+,+,-),(-,+,-) then the z+ face
(reference tools/generate_chexa_mesh.py:93-125); CQUAD4 tubes use
G1=(ix,ic), G2=(ix,ic+1), G3=(ix+1,ic+1), G4=(ix+1,ic)
(reference tools/generate_cquad4_tube.py docstring 18-31).
"""
from __future__ import annotations

import numpy as np

from .model import CHEXA8, CPENTA6, CQUAD4, CTETRA4, LoadCase, Model, Mpc, Pload4


def _hex_grid(nx, ny, nz, h=1.0):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ijk = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], axis=1)
    xyz = ijk.astype(np.float64) * h
    ids = (ijk[:, 0] * (ny + 1) * (nz + 1) + ijk[:, 1] * (nz + 1) + ijk[:, 2] + 1).astype(np.int32)
    return ijk, xyz, ids


def _hex_conn(nx, ny, nz):
    ie, je, ke = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ie, je, ke = ie.reshape(-1), je.reshape(-1), ke.reshape(-1)

    def n(di, dj, dk):
        return (ie + di) * (ny + 1) * (nz + 1) + (je + dj) * (nz + 1) + (ke + dk)

    conn = np.stack([n(0, 0, 0), n(1, 0, 0), n(1, 1, 0), n(0, 1, 0),
                     n(0, 0, 1), n(1, 0, 1), n(1, 1, 1), n(0, 1, 1)], axis=1)
    return conn.astype(np.int32)


def hex_block(nx, ny, nz, h=1.0, E=200e3, nu=0.3, rho=7.85e-9, jitter=0.0, jitter_seed=1,
              E_sigma=0.0, E_seed=2, solid_form=0):
    """Structured CHEXA8 block on [0,nx*h]x[0,ny*h]x[0,nz*h]. Interior nodes get
    uniform jitter of +-jitter*h per axis; E is lognormal with median E when
    E_sigma > 0. The x=0 face is clamped (SPC1 123)."""
    ijk, xyz, ids = _hex_grid(nx, ny, nz, h)
    if jitter > 0.0:
        rng = np.random.default_rng(jitter_seed)
        interior = ((ijk[:, 0] > 0) & (ijk[:, 0] < nx) & (ijk[:, 1] > 0) & (ijk[:, 1] < ny)
                    & (ijk[:, 2] > 0) & (ijk[:, 2] < nz))
        d = rng.uniform(-jitter * h, jitter * h, size=xyz.shape)
        xyz = xyz + d * interior[:, None]
    conn = _hex_conn(nx, ny, nz)
    ne = conn.shape[0]
    if E_sigma > 0.0:
        rng = np.random.default_rng(E_seed)
        Ee = E * np.exp(E_sigma * rng.standard_normal(ne))
    else:
        Ee = np.full(ne, float(E))
    spc_nodes = np.nonzero(ijk[:, 0] == 0)[0].astype(np.int32)
    m = Model(node_ids=ids, xyz=xyz, elem_type=np.full(ne, CHEXA8, np.int32),
              elem_ptr=(np.arange(ne + 1) * 8).astype(np.int32), conn=conn.reshape(-1),
              E=Ee, nu=np.full(ne, float(nu)), rho=np.full(ne, float(rho)),
              solid_form=solid_form, spc_nodes=spc_nodes,
              spc_masks=np.full(len(spc_nodes), 0b000111, np.int32))
    m.meta.update(dict(nx=nx, ny=ny, nz=nz, h=h, ijk=ijk))
    return m


def hex_block_mpc(nx=120, ny=40, nz=40, seed=11, jitter=0.1):
    """MPC workload on the CHEXA8 block (x = 0 clamped): the free end face
    x = nx*h tied to its centre node in T1-T3 (u(g, d) - u(c, d) = 0, the
    face moves as one translation), the y = 0 side face on inclined slides
    a u_x + b u_y = 0 with |a| > |b| drawn per node, and a point FORCE -z at
    the centre node."""
    m = hex_block(nx, ny, nz, jitter=jitter, jitter_seed=seed)
    ijk = m.meta["ijk"]
    end = np.nonzero(ijk[:, 0] == nx)[0]
    c = int(np.nonzero((ijk[:, 0] == nx) & (ijk[:, 1] == ny // 2) & (ijk[:, 2] == nz // 2))[0][0])
    eqs = [[(int(g), d, 1.0), (c, d, -1.0)] for g in end if g != c for d in (1, 2, 3)]
    rng = np.random.default_rng(seed)
    side = np.nonzero((ijk[:, 1] == 0) & (ijk[:, 0] > 0) & (ijk[:, 0] < nx))[0]
    a = rng.uniform(1.0, 2.0, size=len(side))
    b = rng.uniform(-0.9, 0.9, size=len(side))
    eqs += [[(int(g), 1, float(ai)), (int(g), 2, float(bi))] for g, ai, bi in zip(side, a, b)]
    m.mpc = Mpc.from_equations(eqs)
    nodal = np.zeros((m.nn, 6))
    nodal[c, 2] = -1000.0
    m.loads.append(LoadCase(nodal=nodal))
    m.meta.update(dict(name=f"hex{nx}x{ny}x{nz}-mpc-tie-slide", centre=c))
    return m


def config1_hex_cantilever():
    """Plain SOL 101, CPU-runnable: CHEXA8 20x10x10 unit cubes, E=200e3,
    nu=0.3, rho=7.85e-9, SPC123 on x=0, FORCE -1000 in z spread evenly over
    the x=20 face nodes."""
    m = hex_block(20, 10, 10)
    ijk = m.meta["ijk"]
    tip = np.nonzero(ijk[:, 0] == 20)[0]
    nodal = np.zeros((m.nn, 6))
    nodal[tip, 2] = -1000.0 / len(tip)
    m.loads.append(LoadCase(nodal=nodal))
    m.meta["name"] = "hex20x10x10-cantilever"
    return m


def config2_hex_jitter(nx=120, ny=40, nz=40, seed=2024):
    """Mixed-precision workload: CHEXA8 120x40x40 with +-10% interior jitter,
    lognormal E (median 200e3, sigma 0.3), nu=0.3, clamped x=0, uniform
    PLOAD4 p=1.0 on the top (z max) faces."""
    m = hex_block(nx, ny, nz, jitter=0.1, jitter_seed=seed, E_sigma=0.3, E_seed=seed + 1)
    ne = m.ne
    ke = np.arange(ne) % nz
    top = np.nonzero(ke == nz - 1)[0]
    # G1 = corner 5 (local 4), G34 = corner 7 (local 6): the z+ face.
    pl = Pload4(elem=top, g1=np.full(len(top), 4), g34=np.full(len(top), 6),
                p=np.ones((len(top), 4)))
    m.loads.append(LoadCase(nodal=np.zeros((m.nn, 6)), pload4=pl))
    m.meta["name"] = f"hex{nx}x{ny}x{nz}-jitter-lognormalE-pload4"
    return m


def config4_simp(nx=80, ny=40, nz=20, seed=7, E0=200e3, Emin=1e-3):
    """Adjoint workload: CHEXA8 80x40x20, rho_e ~ U[0.01,1], SIMP modulus
    E_e = Emin + rho_e^3 (E0 - Emin), nu=0.3, clamped x=0, point FORCE -z at
    the mid-edge node (nx, ny/2, 0) of the free end."""
    m = hex_block(nx, ny, nz)
    rng = np.random.default_rng(seed)
    dens = rng.uniform(0.01, 1.0, size=m.ne)
    m.E = Emin + dens ** 3 * (E0 - Emin)
    ijk = m.meta["ijk"]
    tip = np.nonzero((ijk[:, 0] == nx) & (ijk[:, 1] == ny // 2) & (ijk[:, 2] == 0))[0]
    nodal = np.zeros((m.nn, 6))
    nodal[tip, 2] = -1000.0
    m.loads.append(LoadCase(nodal=nodal))
    m.meta.update(dict(name=f"hex{nx}x{ny}x{nz}-simp", density=dens, E0=E0, Emin=Emin,
                       tip_node=int(tip[0])))
    return m


def cquad4_tube(n_circ=400, n_axial=800, radius=50.0, length=400.0, t=2.0, E=70e3, nu=0.33,
                rho=2.7e-9, n_cases=32, seed=11, load_scale=1e3):
    """Batched-RHS workload: CQUAD4 cylindrical tube along x, node id =
    ix*n_circ + ic + 1, one end (x=0) clamped 123456, n_cases load cases each
    a seeded resultant force+moment in U[-1,1]*load_scale spread evenly over
    the free-end ring nodes."""
    ix, ic = np.meshgrid(np.arange(n_axial + 1), np.arange(n_circ), indexing="ij")
    ix, ic = ix.reshape(-1), ic.reshape(-1)
    th = 2.0 * np.pi * ic / n_circ
    xyz = np.stack([length * ix / n_axial, radius * np.cos(th), radius * np.sin(th)], axis=1)
    ids = (ix * n_circ + ic + 1).astype(np.int32)
    ex, ec = np.meshgrid(np.arange(n_axial), np.arange(n_circ), indexing="ij")
    ex, ec = ex.reshape(-1), ec.reshape(-1)

    def node(a, c):
        return a * n_circ + (c % n_circ)

    conn = np.stack([node(ex, ec), node(ex, ec + 1), node(ex + 1, ec + 1), node(ex + 1, ec)], axis=1)
    ne = conn.shape[0]
    spc_nodes = np.nonzero(ix == 0)[0].astype(np.int32)
    m = Model(node_ids=ids, xyz=xyz, elem_type=np.full(ne, CQUAD4, np.int32),
              elem_ptr=(np.arange(ne + 1) * 4).astype(np.int32), conn=conn.reshape(-1).astype(np.int32),
              E=np.full(ne, float(E)), nu=np.full(ne, float(nu)), rho=np.full(ne, float(rho)),
              thickness=np.full(ne, float(t)), spc_nodes=spc_nodes,
              spc_masks=np.full(len(spc_nodes), 0b111111, np.int32))
    tip = np.nonzero(ix == n_axial)[0]
    rng = np.random.default_rng(seed)
    for _ in range(n_cases):
        fm = rng.uniform(-1.0, 1.0, size=6) * load_scale
        nodal = np.zeros((m.nn, 6))
        nodal[tip, :] = fm / len(tip)
        m.loads.append(LoadCase(nodal=nodal))
    m.meta.update(dict(name=f"cquad4-tube-{n_circ}x{n_axial}", n_circ=n_circ, n_axial=n_axial))
    return m


_TET_A = np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]])
_TET_B = np.array([[0, 1, 2, 5], [0, 2, 3, 7], [0, 5, 7, 4], [2, 5, 6, 7], [0, 2, 5, 7]])
_WEDGES = np.array([[0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7]])


def mixed_block(nx=120, ny=60, nz=54, jitter=0.1, seed=5, n_cases=8, hex_frac=0.4,
                wedge_frac=0.1, n_force_nodes=16, force_scale=1e3):
    """Mixed-mesh workload: a jittered nx x ny x nz cell block (h = 1) whose
    cells are CHEXA8 for i < hex_frac*nx, two CPENTA6 each for the next
    wedge_frac*nx layers, and five CTETRA4 each beyond (alternating split for
    conformity). Every element takes one of two materials by a seeded draw
    (E 200e3 / nu 0.3 / rho 7.85e-9, or 70e3 / 0.33 / 2.7e-9). The base z = 0
    is clamped (123). Load case j: GRAV (0, 0, -9810) plus FORCE components
    U[-1,1]*force_scale on n_force_nodes seeded top-surface nodes."""
    ijk, xyz, ids = _hex_grid(nx, ny, nz)
    rng = np.random.default_rng(seed)
    interior = ((ijk[:, 0] > 0) & (ijk[:, 0] < nx) & (ijk[:, 1] > 0) & (ijk[:, 1] < ny)
                & (ijk[:, 2] > 0) & (ijk[:, 2] < nz))
    xyz = xyz + rng.uniform(-jitter, jitter, size=xyz.shape) * interior[:, None]
    cells = _hex_conn(nx, ny, nz)                        # [nc,8], x-major
    ci = np.arange(len(cells)) // (ny * nz)             # cell i index
    cjk = np.arange(len(cells)) % (ny * nz)
    parity = (ci + cjk // nz + cjk % nz) % 2
    i_hex, i_wedge = int(round(hex_frac * nx)), int(round((hex_frac + wedge_frac) * nx))
    hexes = cells[ci < i_hex]
    wsel = cells[(ci >= i_hex) & (ci < i_wedge)]
    wedges = wsel[:, _WEDGES].reshape(-1, 6)
    tsel = (ci >= i_wedge)
    tets = np.concatenate([cells[tsel & (parity == 0)][:, _TET_A].reshape(-1, 4),
                           cells[tsel & (parity == 1)][:, _TET_B].reshape(-1, 4)])
    X = xyz[tets]
    neg = np.einsum("ij,ij->i", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    et = np.concatenate([np.full(len(hexes), CHEXA8), np.full(len(wedges), CPENTA6),
                         np.full(len(tets), CTETRA4)]).astype(np.int32)
    conn = np.concatenate([hexes.reshape(-1), wedges.reshape(-1), tets.reshape(-1)]).astype(np.int32)
    npe = np.where(et == CHEXA8, 8, np.where(et == CPENTA6, 6, 4))
    ep = np.concatenate([[0], np.cumsum(npe)]).astype(np.int32)
    ne = len(et)
    mat = rng.random(ne) < 0.5
    spc_nodes = np.nonzero(ijk[:, 2] == 0)[0].astype(np.int32)
    m = Model(node_ids=ids, xyz=xyz, elem_type=et, elem_ptr=ep, conn=conn,
              E=np.where(mat, 200e3, 70e3), nu=np.where(mat, 0.3, 0.33),
              rho=np.where(mat, 7.85e-9, 2.7e-9), spc_nodes=spc_nodes,
              spc_masks=np.full(len(spc_nodes), 0b000111, np.int32))
    top = np.nonzero(ijk[:, 2] == nz)[0]
    for j in range(n_cases):
        r = np.random.default_rng(seed * 1000 + j)
        nodal = np.zeros((m.nn, 6))
        pick = r.choice(top, n_force_nodes, replace=False)
        nodal[pick, 0:3] = r.uniform(-1.0, 1.0, (n_force_nodes, 3)) * force_scale
        m.loads.append(LoadCase(nodal=nodal, grav=np.array([0.0, 0.0, -9810.0])))
    m.meta.update(dict(name=f"mixed{nx}x{ny}x{nz}-hex-penta-tet", nx=nx, ny=ny, nz=nz, ijk=ijk,
                       counts=dict(CHEXA8=len(hexes), CPENTA6=len(wedges), CTETRA4=len(tets))))
    return m


def bdf_text(m: Model, fmt: str = "free") -> str:
    """The model as BDF text: SOL 101 case control with one SUBCASE per load
    case (LOAD = k+1, SPC = 1, DISPLACEMENT/STRESS = ALL), one MAT1 + PSOLID (or
    PSHELL) per element so per-element E survives, SPC1 (THRU for runs of
    consecutive IDs), FORCE/MOMENT, PLOAD4 with its face grids, GRAV.
    fmt "free": every card comma-separated; "mixed": GRID in large field
    (GRID* + '*' continuation), elements and PLOAD4 in small field (8-column,
    blank-label continuations), the rest free field. Values are written with
    repr (free field) or must fit the fixed fields exactly (the caller rounds
    coordinates for "mixed")."""
    ids = np.asarray(m.node_ids)
    eids = m.element_ids()
    out = ["SOL 101", "CEND"]
    for k in range(max(1, len(m.loads))):
        out += [f"SUBCASE {m.subcase_id(k)}", f"  LOAD = {k + 1}", "  SPC = 1",
                "  DISPLACEMENT = ALL", "  STRESS = ALL"]
    out.append("BEGIN BULK")
    if m.shell_form == 1:
        out.append("PARAM,SHELLFORM,MINDLIN")

    def small(*fields):
        return "".join(f"{str(x):<8s}" for x in fields).rstrip()

    for i, g in enumerate(ids):
        x, y, z = (repr(float(v)) for v in m.xyz[i])
        if fmt == "mixed":
            out.append(f"{'GRID*':<8s}{g:<16d}{'':16s}{float(m.xyz[i, 0]):<16.9E}"
                       f"{float(m.xyz[i, 1]):<16.9E}".rstrip())
            out.append(f"{'*':<8s}{float(m.xyz[i, 2]):<16.9E}".rstrip())
        else:
            out.append(f"GRID,{g},,{x},{y},{z}")
    names = {CHEXA8: "CHEXA", CPENTA6: "CPENTA", CTETRA4: "CTETRA", CQUAD4: "CQUAD4"}
    for e in range(m.ne):
        t = int(m.elem_type[e])
        nodes = [int(ids[c]) for c in m.conn[m.elem_ptr[e]:m.elem_ptr[e + 1]]]
        eid = int(eids[e])
        if fmt == "mixed":
            head = [names[t], eid, eid] + nodes[:6]
            out.append(small(*head))
            if len(nodes) > 6:
                out.append(small("", *nodes[6:]))
        else:
            line = f"{names[t]},{eid},{eid}," + ",".join(map(str, nodes[:6]))
            if len(nodes) > 6:
                out.append(line + ",+")
                out.append("+," + ",".join(map(str, nodes[6:])))
            else:
                out.append(line)
        rho = repr(float(m.rho[e])) if m.rho is not None else ""
        out.append(f"MAT1,{eid},{repr(float(m.E[e]))},,{repr(float(m.nu[e]))},{rho}")
        if t == CQUAD4:
            out.append(f"PSHELL,{eid},{eid},{repr(float(m.thickness[e]))}")
        else:
            out.append(f"PSOLID,{eid},{eid}" + (",,,,SRI" if m.solid_form == 1 else ""))
    # SPC1 per mask, THRU for runs of consecutive IDs
    for mask in sorted(set(int(x) for x in m.spc_masks)):
        comp = "".join(str(d + 1) for d in range(6) if mask >> d & 1)
        g = np.sort(ids[m.spc_nodes[m.spc_masks == mask]])
        runs = np.split(g, np.nonzero(np.diff(g) != 1)[0] + 1)
        for r in runs:
            if len(r) >= 3:
                out.append(f"SPC1,1,{comp},{r[0]},THRU,{r[-1]}")
            else:
                for v in r:
                    out.append(f"SPC1,1,{comp},{v}")
    for k, lc in enumerate(m.loads):
        sid = k + 1
        if lc.nodal is not None:
            for i in np.nonzero(np.any(lc.nodal != 0.0, axis=1))[0]:
                for name, off in (("FORCE", 0), ("MOMENT", 3)):
                    v = lc.nodal[i, off:off + 3]
                    if np.any(v != 0.0):
                        out.append(f"{name},{sid},{ids[i]},0,1.0," +
                                   ",".join(repr(float(a)) for a in v))
        if lc.pload4 is not None:
            pl = lc.pload4
            for j, e in enumerate(pl.elem):
                c = m.conn[m.elem_ptr[e]:m.elem_ptr[e + 1]]
                g1 = "" if pl.g1[j] < 0 else int(ids[c[pl.g1[j]]])
                g34 = "" if pl.g34[j] < 0 else int(ids[c[pl.g34[j]]])
                ps = [repr(float(p)) for p in pl.p[j]]
                if fmt == "mixed":
                    out.append(small("PLOAD4", sid, int(eids[e]), *[f"{float(p):.6g}" for p in pl.p[j]],
                                     g1, g34))
                else:
                    out.append(f"PLOAD4,{sid},{int(eids[e])}," + ",".join(ps) + f",{g1},{g34}")
        if lc.grav is not None:
            out.append(f"GRAV,{sid},0,1.0," + ",".join(repr(float(a)) for a in lc.grav))
    out.append("ENDDATA")
    return "\n".join(out) + "\n"
