"""Native BDF input (libvbh's vbh_deck_* reader) into a ``Model``.

The reader restates the reference's BdfParser (src/io/bdf_parser.cpp) for the
cards of the linear-static path and parses large decks in parallel chunks; it
needs no GPU. ``read_deck`` returns the same ``Model`` layout as the
pure-Python ``bdf.read_bdf`` (nodes in ascending ID order, elements in file
order, one ``LoadCase`` per SUBCASE)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .backend import _check
from .model import LoadCase, Model, Mpc, Pload4


def _read(handle) -> Model:
    lib = L.load()
    sz = L.vbh_deck_sizes()
    _check(lib.vbh_deck_info(handle, C.byref(sz)))
    nn, ne, nc = sz.nnodes, sz.nelem, sz.nconn
    ids = np.empty(nn, np.int32)
    xyz = np.empty((nn, 3))
    eids = np.empty(ne, np.int32)
    et = np.empty(ne, np.int32)
    ep = np.empty(ne + 1, np.int32)
    conn = np.empty(nc, np.int32)
    E, nu, rho, th = (np.empty(ne) for _ in range(4))
    _check(lib.vbh_deck_mesh(handle, L.ip(ids), L.dp(xyz), L.ip(eids), L.ip(et), L.ip(ep),
                             L.ip(conn), L.dp(E), L.dp(nu), L.dp(rho), L.dp(th)))
    from .model import CQUAD4
    m = Model(node_ids=ids, xyz=xyz, elem_type=et, elem_ptr=ep, conn=conn, E=E, nu=nu, rho=rho,
              thickness=th if np.any(et == CQUAD4) else None, solid_form=sz.solid_form,
              shell_form=sz.shell_form, elem_ids=eids)
    sub_ids = []
    for k in range(sz.nsubcase):
        si = L.vbh_deck_subcase_info()
        _check(lib.vbh_deck_subcase(handle, k, C.byref(si)))
        sn = np.empty(si.nspc, np.int32)
        sm = np.empty(si.nspc, np.int32)
        nodal = np.empty((nn, 6))
        pe = np.empty(si.npload4, np.int32)
        g1 = np.empty(si.npload4, np.int32)
        g34 = np.empty(si.npload4, np.int32)
        pp = np.empty((si.npload4, 4))
        _check(lib.vbh_deck_subcase_data(handle, k, L.ip(sn), L.ip(sm), L.dp(nodal), L.ip(pe),
                                         L.ip(g1), L.ip(g34), L.dp(pp)))
        if k == 0:  # the model's SPC and MPC sets are the first subcase's (one K per deck)
            m.spc_nodes, m.spc_masks = sn, sm
            if si.nmpc:
                mp = Mpc(ptr=np.empty(si.nmpc + 1, np.int32), node=np.empty(si.nmpc_terms, np.int32),
                         dof=np.empty(si.nmpc_terms, np.int32), coeff=np.empty(si.nmpc_terms))
                _check(lib.vbh_deck_subcase_mpc(handle, k, L.ip(mp.ptr), L.ip(mp.node),
                                                L.ip(mp.dof), L.dp(mp.coeff)))
                m.mpc = mp
        pl = None
        if si.npload4:
            # face grids as element-local corner indices (-1 = default face)
            def local(e, g):
                if g < 0:
                    return -1
                c = conn[ep[e]:ep[e + 1]]
                return int(np.nonzero(c == g)[0][0])
            pl = Pload4(elem=pe.astype(np.int64),
                        g1=np.array([local(e, g) for e, g in zip(pe, g1)]),
                        g34=np.array([local(e, g) for e, g in zip(pe, g34)]), p=pp)
        grav = np.array(si.grav[:]) if any(si.grav[:]) else None
        m.loads.append(LoadCase(nodal=nodal, pload4=pl, grav=grav))
        sub_ids.append(si.id)
    m.subcase_ids = sub_ids
    if sz.nrigid_mpc:
        act = np.empty(nn, np.int32)
        rg = Mpc(ptr=np.empty(sz.nrigid_mpc + 1, np.int32), node=np.empty(sz.nrigid_terms, np.int32),
                 dof=np.empty(sz.nrigid_terms, np.int32), coeff=np.empty(sz.nrigid_terms))
        _check(lib.vbh_deck_rigid(handle, L.ip(act), L.ip(rg.ptr), L.ip(rg.node), L.ip(rg.dof),
                                  L.dp(rg.coeff)))
        m.rigid_mpc, m.rigid_active = rg, act
    cd = np.empty(nn, np.int32)
    rot = np.empty((nn, 3, 3))
    _check(lib.vbh_deck_frames(handle, L.ip(cd), L.dp(rot)))
    if np.any(cd != 0):
        m.cd, m.cd_rot = cd, rot
    m.meta["unsupported_cards"] = [lib.vbh_deck_unsupported(handle, i).decode()
                                   for i in range(sz.nunsupported)]
    return m


def read_deck(path_or_text: str, is_text: bool = False) -> Model:
    """Parse a BDF file (or its text) with the native reader."""
    lib = L.load()
    h = C.c_void_p()
    if is_text:
        raw = path_or_text.encode()
        _check(lib.vbh_deck_parse(raw, len(raw), C.byref(h)))
    else:
        _check(lib.vbh_deck_read(str(path_or_text).encode(), C.byref(h)))
    try:
        return _read(h)
    finally:
        lib.vbh_deck_free(h)
