"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
package ``pytential_amd`` never imports it.

The oracle restates the reference algorithm (see ``qbx_oracle.h`` for the
file:line map).  Parity status of each function is recorded in its test:
closed forms, brute force and scipy pin the expansions, the lists and the
association; value-level parity with sumpy/boxtree is unpinnable (those
packages are not vendored and cannot be imported here).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc_qbx.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libhelmholtz_utils.so")

_lib = None


def build(force: bool = False, with_ref: bool = True) -> None:
    args = ["make", "-s", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    if with_ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        _lib = ct.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ct.c_void_p
IP = ct.POINTER(ct.c_int)


class OrcTree(ct.Structure):
    _fields_ = [
        ("dim", ct.c_int), ("nsources", ct.c_int), ("ntargets", ct.c_int),
        ("nboxes", ct.c_int), ("nlevels", ct.c_int), ("maxl", ct.c_int),
        ("root_min", ct.c_double * 3), ("root_extent", ct.c_double),
        ("box_levels", IP), ("box_parent_ids", IP), ("box_child_ids", IP),
        ("box_coords", IP), ("box_centers", ct.POINTER(ct.c_double)),
        ("box_source_starts", IP), ("box_source_counts_nonchild", IP),
        ("box_source_counts_cumul", IP),
        ("box_target_starts", IP), ("box_target_counts_nonchild", IP),
        ("box_target_counts_cumul", IP),
        ("level_start_box_nrs", IP), ("user_source_ids", IP),
        ("sorted_target_ids", IP),
    ]


class OrcTrav(ct.Structure):
    _fields_ = [
        ("nlevels", ct.c_int),
        ("nsource_boxes", ct.c_int), ("nsource_parent_boxes", ct.c_int),
        ("ntarget_boxes", ct.c_int), ("nttp_boxes", ct.c_int),
        ("source_boxes", IP), ("level_start_source_box_nrs", IP),
        ("source_parent_boxes", IP), ("level_start_source_parent_box_nrs", IP),
        ("target_boxes", IP), ("level_start_target_box_nrs", IP),
        ("ttp_boxes", IP), ("level_start_ttp_box_nrs", IP),
        ("n1", ct.c_int), ("n2", ct.c_int), ("n3", ct.c_int), ("n4", ct.c_int),
        ("l1_starts", IP), ("l1_lists", IP), ("l2_starts", IP), ("l2_lists", IP),
        ("l3_starts", IP), ("l3_lists", IP), ("l4_starts", IP), ("l4_lists", IP),
    ]


class KParams(ct.Structure):
    _fields_ = [("kernel", ct.c_int), ("k", ct.c_double), ("source_kind", ct.c_int),
                ("qbx_order", ct.c_int), ("level_orders", IP), ("nthreads", ct.c_int),
                ("target_deriv", ct.c_int), ("target_dir", P)]


def _declare(L):
    L.orc_tree_build.argtypes = [ct.c_int, ct.c_int, P, ct.c_int, P, P, P, P,
                                 ct.c_int, ct.c_double, ct.c_int, ct.POINTER(OrcTree)]
    L.orc_tree_free.argtypes = [ct.POINTER(OrcTree)]
    L.orc_traverse.argtypes = [ct.POINTER(OrcTree), ct.c_int, ct.c_int, ct.c_int,
                               ct.POINTER(OrcTrav)]
    L.orc_trav_free.argtypes = [ct.POINTER(OrcTrav)]
    L.orc_target_assoc.argtypes = [ct.c_int, ct.c_int, P, P, ct.c_int, P, P,
                                   ct.c_int, P, P, ct.c_int, P, P]
    L.orc_refine_flags_for_failed.argtypes = [ct.c_int, ct.c_int, P, P, ct.c_int, P,
                                              ct.c_int, P, P, P]
    L.orc_qbx_fmm.argtypes = [ct.POINTER(OrcTree), ct.POINTER(OrcTrav),
                              ct.POINTER(KParams), P, P, P, ct.c_int, P, P, P, P, P, P]
    L.orc_qbx_fmm_ex.argtypes = [ct.POINTER(OrcTree), ct.POINTER(OrcTrav),
                                 ct.POINTER(KParams), P, P, P, ct.c_int, P, P, P, P, P, P,
                                 P, P, P]
    L.orc_direct_qbx.argtypes = [ct.c_int, ct.POINTER(KParams), ct.c_int, P, P, P,
                                 ct.c_int, P, P, ct.c_int, P, P, P]
    L.orc_p2e.argtypes = [ct.POINTER(KParams), ct.c_int, ct.c_int, P, ct.c_double,
                          ct.c_int, P, P, P, P]
    L.orc_e2p.argtypes = [ct.POINTER(KParams), ct.c_int, ct.c_int, P, ct.c_double,
                          P, ct.c_int, P, P]
    L.orc_e2e.argtypes = [ct.POINTER(KParams), ct.c_int, ct.c_int, ct.c_int,
                          P, ct.c_double, P, ct.c_double, P, P]
    L.orc_p2p.argtypes = [ct.POINTER(KParams), ct.c_int, P, P, P, ct.c_int, P, P]
    L.orc_bessel_jy.argtypes = [ct.c_int, ct.c_double, P, P]
    L.orc_sh3d.argtypes = [ct.c_int, P, P, P]
    L.orc_gaunt3.argtypes = [ct.c_int] * 5
    L.orc_gaunt3.restype = ct.c_double
    L.orc_tsqbx.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_int,
                            ct.c_int, P, ct.c_int, P, ct.c_int, P,
                            ct.c_int, P, P, P, P, P, P, P, P, P, P, P, P]
    L.orc_sph_jn.argtypes = [ct.c_int, ct.c_double, ct.c_double, ct.c_double, P, P]
    L.orc_sph_hn.argtypes = [ct.c_int, ct.c_double, ct.c_double, ct.c_double, P, P]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _arr(ptr, n, dtype=np.int32):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Tree:
    """Oracle tree (boxtree ``Tree`` field names)."""

    def __init__(self, sources, targets, target_radii=None, source_weights=None,
                 target_weights=None, max_leaf_refine_weight=64,
                 stick_out_factor=0.5, extent_norm="l2"):
        L = lib()
        self._s = OrcTree()
        src = _f64(sources)
        tgt = _f64(targets)
        dim = src.shape[0]
        self._keep = [src, tgt]
        rc = L.orc_tree_build(
            dim, src.shape[1], _ptr(src), tgt.shape[1], _ptr(tgt),
            _ptr(_f64(target_radii)), _ptr(_i32(source_weights)),
            _ptr(_i32(target_weights)), max_leaf_refine_weight,
            stick_out_factor, int(extent_norm == "linf"), ct.byref(self._s))
        assert rc == 0
        s = self._s
        nb = s.nboxes
        nchild = 2 ** dim
        self.dimensions = dim
        self.nsources, self.ntargets, self.nboxes, self.nlevels = (
            s.nsources, s.ntargets, s.nboxes, s.nlevels)
        self.root_extent = s.root_extent
        self.bounding_box_min = np.array(s.root_min[:dim])
        self.box_levels = _arr(s.box_levels, nb)
        self.box_parent_ids = _arr(s.box_parent_ids, nb)
        self.box_child_ids = _arr(s.box_child_ids, nb * nchild).reshape(nchild, nb)
        self.box_coords = _arr(s.box_coords, nb * dim).reshape(dim, nb)
        self.box_centers = _arr(s.box_centers, nb * dim, np.float64).reshape(dim, nb)
        for k in ["box_source_starts", "box_source_counts_nonchild",
                  "box_source_counts_cumul", "box_target_starts",
                  "box_target_counts_nonchild", "box_target_counts_cumul"]:
            setattr(self, k, _arr(getattr(s, k), nb))
        self.level_start_box_nrs = _arr(s.level_start_box_nrs, s.nlevels + 1)
        self.user_source_ids = _arr(s.user_source_ids, s.nsources)
        self.sorted_target_ids = _arr(s.sorted_target_ids, s.ntargets)
        self.sources = src[:, self.user_source_ids]
        self.targets = tgt[:, self.sorted_target_ids]

    def __del__(self):
        try:
            lib().orc_tree_free(ct.byref(self._s))
        except Exception:
            pass


class Traversal:
    def __init__(self, tree: Tree, well_sep_is_n_away=2, targets_have_extent=True,
                 from_sep_smaller_min_nsources_cumul=15):
        L = lib()
        self.tree = tree
        self._s = OrcTrav()
        rc = L.orc_traverse(ct.byref(tree._s), well_sep_is_n_away,
                            int(targets_have_extent),
                            from_sep_smaller_min_nsources_cumul, ct.byref(self._s))
        assert rc == 0
        s = self._s
        nl = s.nlevels
        self.source_boxes = _arr(s.source_boxes, s.nsource_boxes)
        self.source_parent_boxes = _arr(s.source_parent_boxes, s.nsource_parent_boxes)
        self.target_boxes = _arr(s.target_boxes, s.ntarget_boxes)
        self.target_or_target_parent_boxes = _arr(s.ttp_boxes, s.nttp_boxes)
        self.level_start_source_box_nrs = _arr(s.level_start_source_box_nrs, nl + 1)
        self.level_start_source_parent_box_nrs = _arr(
            s.level_start_source_parent_box_nrs, nl + 1)
        self.level_start_target_box_nrs = _arr(s.level_start_target_box_nrs, nl + 1)
        self.level_start_target_or_target_parent_box_nrs = _arr(
            s.level_start_ttp_box_nrs, nl + 1)
        nt, nttp = s.ntarget_boxes, s.nttp_boxes
        self.neighbor_source_boxes_starts = _arr(s.l1_starts, nt + 1)
        self.neighbor_source_boxes_lists = _arr(s.l1_lists, s.n1)
        self.from_sep_siblings_starts = _arr(s.l2_starts, nttp + 1)
        self.from_sep_siblings_lists = _arr(s.l2_lists, s.n2)
        self.from_sep_smaller_starts = _arr(s.l3_starts, nt + 1)
        self.from_sep_smaller_lists = _arr(s.l3_lists, s.n3)
        self.from_sep_bigger_starts = _arr(s.l4_starts, nttp + 1)
        self.from_sep_bigger_lists = _arr(s.l4_lists, s.n4)

    def __del__(self):
        try:
            lib().orc_trav_free(ct.byref(self._s))
        except Exception:
            pass


def target_assoc(src1, tunnel, centers, radii_with_tol, targets, flags,
                 accelerated=True):
    L = lib()
    src1, tunnel, centers, rtol, targets = map(
        _f64, (src1, tunnel, centers, radii_with_tol, targets))
    flags = _i32(flags)
    nt = targets.shape[1]
    ttc = np.empty(nt, np.int32)
    status = np.empty(nt, np.int32)
    L.orc_target_assoc(src1.shape[0], src1.shape[1], _ptr(src1), _ptr(tunnel),
                       centers.shape[1], _ptr(centers), _ptr(rtol),
                       nt, _ptr(targets), _ptr(flags), int(accelerated),
                       _ptr(ttc), _ptr(status))
    return ttc, status


def refine_flags_for_failed(src1, tunnel, el_starts, targets, status):
    L = lib()
    src1, tunnel, targets = map(_f64, (src1, tunnel, targets))
    el_starts, status = _i32(el_starts), _i32(status)
    nel = len(el_starts) - 1
    out = np.empty(nel, np.int32)
    L.orc_refine_flags_for_failed(src1.shape[0], src1.shape[1], _ptr(src1),
                                  _ptr(tunnel), nel, _ptr(el_starts),
                                  targets.shape[1], _ptr(targets), _ptr(status),
                                  _ptr(out))
    return out


KERNEL_CODES = {("laplace", 2): 0, ("helmholtz", 2): 1, ("laplace", 3): 2,
                ("helmholtz", 3): 3}


def kparams(kernel="laplace", k=0.0, source_kind="S", qbx_order=4,
            level_orders=None, nthreads=0, dim=2, target_dir=None):
    """``target_dir`` (dim, n), indexed like the targets of the call: evaluate
    the derivative along it instead of the potential (2D)."""
    lo = _i32(level_orders if level_orders is not None else [0])
    td = None if target_dir is None else _f64(target_dir)
    kp = KParams(KERNEL_CODES[(kernel, dim)], float(k),
                 0 if source_kind == "S" else 1, qbx_order,
                 lo.ctypes.data_as(IP), nthreads, int(td is not None), _ptr(td))
    kp._lo = lo
    kp._td = td
    return kp


def qbx_fmm(tree: Tree, trav: Traversal, kp, sources, normals, strengths,
            centers, radii, all_targets, user_target_to_center):
    """All arrays in user order; ``all_targets`` = [centers, user targets]."""
    L = lib()
    src, nrm, ctr, rad, tgt = map(_f64, (sources, normals, centers, radii, all_targets))
    st = _c128(strengths)
    ttc = _i32(user_target_to_center)
    out = np.zeros(tgt.shape[1], np.complex128)
    times = np.zeros(16)
    rc = L.orc_qbx_fmm(ct.byref(tree._s), ct.byref(trav._s), ct.byref(kp),
                       _ptr(src), _ptr(nrm), _ptr(st), ctr.shape[1], _ptr(ctr),
                       _ptr(rad), _ptr(tgt), _ptr(ttc), _ptr(out), _ptr(times))
    assert rc == 0
    return out, times


def direct_qbx(kp, sources, normals, strengths, centers, radii, targets, ttc):
    L = lib()
    src, nrm, ctr, rad, tgt = map(_f64, (sources, normals, centers, radii, targets))
    st = _c128(strengths)
    ttc = _i32(ttc)
    out = np.zeros(tgt.shape[1], np.complex128)
    rc = L.orc_direct_qbx(src.shape[0], ct.byref(kp), src.shape[1], _ptr(src),
                          _ptr(nrm), _ptr(st), ctr.shape[1], _ptr(ctr), _ptr(rad),
                          tgt.shape[1], _ptr(tgt), _ptr(ttc), _ptr(out))
    assert rc == 0
    return out


def ncoeff(kernel, p, dim=2):
    if dim == 3:
        return (p + 1) ** 2
    return p + 1 if kernel == "laplace" else 2 * p + 1


def p2e(kp, kernel, local, order, center, rscale, sources, normals, strengths):
    L = lib()
    src, nrm = _f64(sources), _f64(normals)
    st = _c128(strengths)
    c = _f64(center)
    out = np.zeros(ncoeff(kernel, order, len(c)), np.complex128)
    L.orc_p2e(ct.byref(kp), int(local), order, _ptr(c), rscale, src.shape[1],
              _ptr(src), _ptr(nrm), _ptr(st), _ptr(out))
    return out


def e2p(kp, local, order, center, rscale, coeffs, targets):
    L = lib()
    tgt = _f64(targets)
    cf = _c128(coeffs)
    c = _f64(center)
    out = np.zeros(tgt.shape[1], np.complex128)
    L.orc_e2p(ct.byref(kp), int(local), order, _ptr(c), rscale, _ptr(cf),
              tgt.shape[1], _ptr(tgt), _ptr(out))
    return out


E2E_KIND = {"m2m": 0, "m2l": 1, "l2l": 2}


def e2e(kp, kernel, kind, p1, p2, c1, r1, c2, r2, coeffs):
    L = lib()
    cf = _c128(coeffs)
    out = np.zeros(ncoeff(kernel, p2, len(np.ravel(c1))), np.complex128)
    L.orc_e2e(ct.byref(kp), E2E_KIND[kind], p1, p2, _ptr(_f64(c1)), r1,
              _ptr(_f64(c2)), r2, _ptr(cf), _ptr(out))
    return out


def gaunt3(n, m, nu, mu, l):
    """int Y_n^m conj(Y_nu^mu) conj(Y_l^(m-mu)) over the unit sphere
    (orthonormal harmonics, Condon-Shortley phase; fmmh3d_oracle.c)."""
    return lib().orc_gaunt3(n, m, nu, mu, l)


def p2p(kp, sources, normals, strengths, targets):
    L = lib()
    src, nrm, tgt = _f64(sources), _f64(normals), _f64(targets)
    st = _c128(strengths)
    out = np.zeros(tgt.shape[1], np.complex128)
    L.orc_p2p(ct.byref(kp), src.shape[1], _ptr(src), _ptr(nrm), _ptr(st),
              tgt.shape[1], _ptr(tgt), _ptr(out))
    return out


class GeometryData:
    """Oracle counterpart of ``QBXFMMGeometryData`` (qbx/geometry.py:361-1040)
    for one source curve and one list of user targets."""

    def __init__(self, discr, targets, target_sides, *, max_leaf_refine_weight=64,
                 stick_out_factor=0.5, well_sep_is_n_away=2,
                 from_sep_smaller_min_nsources_cumul=None,
                 target_association_tolerance=1.0e3 * 2.220446049250313e-16):
        import time
        t0 = time.perf_counter()
        self.discr = discr
        self.centers = discr.centers()
        self.radii = discr.center_radii()
        self.ncenters = self.centers.shape[1]
        self.user_targets = np.ascontiguousarray(targets, np.float64)
        ttc, status = target_assoc(
            discr.nodes, discr.tunnel_radii(), self.centers,
            self.radii * (1 + target_association_tolerance),
            self.user_targets, target_sides)
        self.target_status = status
        self.target_to_center = ttc
        if np.any(status == 1):
            self.refine_flags = refine_flags_for_failed(
                discr.nodes, discr.tunnel_radii(), discr.element_to_source_starts(),
                self.user_targets, status)
        self.user_target_to_center = np.concatenate(
            [np.full(self.ncenters, -1, np.int32), ttc]).astype(np.int32)
        self.all_targets = np.concatenate([self.centers, self.user_targets], axis=1)
        radii_t = np.concatenate([self.radii, np.zeros(self.user_targets.shape[1])])
        tw = np.concatenate([np.ones(self.ncenters, np.int32),
                             (ttc == -1).astype(np.int32)])
        t1 = time.perf_counter()
        self.tree = Tree(discr.quad_nodes, self.all_targets, radii_t,
                         None, tw, max_leaf_refine_weight, stick_out_factor)
        if from_sep_smaller_min_nsources_cumul is None:
            # qbx/__init__.py:303-310: 15 for curves, 30 for surfaces
            from_sep_smaller_min_nsources_cumul = 15 if discr.nodes.shape[0] == 2 else 30
        self.from_sep_smaller_min_nsources_cumul = from_sep_smaller_min_nsources_cumul
        self.trav = Traversal(self.tree, well_sep_is_n_away, True,
                              from_sep_smaller_min_nsources_cumul)
        t2 = time.perf_counter()
        self.setup_times = {"target_assoc": t1 - t0, "tree+traversal": t2 - t1}

    def fmm(self, kernel, k, source_kind, qbx_order, level_orders, strengths,
            nthreads=0, return_stages=False, target_dir=None):
        """``target_dir`` (dim, n_user_targets): the derivative along it at
        each user target instead of the potential."""
        dim = self.discr.nodes.shape[0]
        td = None
        if target_dir is not None:
            td = np.concatenate([np.zeros((dim, self.ncenters)),
                                 np.asarray(target_dir, np.float64)], axis=1)
        kp = kparams(kernel, k, source_kind, qbx_order, level_orders, nthreads, dim, td)
        if not return_stages:
            out, times = qbx_fmm(self.tree, self.trav, kp, self.discr.quad_nodes,
                                 self.discr.quad_normals, strengths, self.centers,
                                 self.radii, self.all_targets, self.user_target_to_center)
            return out[self.ncenters:], times
        pmax = max(level_orders[:self.tree.nlevels])
        ncf, ncq = ncoeff(kernel, pmax, dim), ncoeff(kernel, qbx_order, dim)
        mp = np.zeros((self.tree.nboxes, ncf), np.complex128)
        lc = np.zeros((self.tree.nboxes, ncf), np.complex128)
        qx = np.zeros((self.ncenters, ncq), np.complex128)
        src, nrm, ctr, rad, tgt = map(_f64, (self.discr.quad_nodes, self.discr.quad_normals,
                                            self.centers, self.radii, self.all_targets))
        st = _c128(strengths)
        ttc = _i32(self.user_target_to_center)
        out = np.zeros(tgt.shape[1], np.complex128)
        times = np.zeros(16)
        rc = lib().orc_qbx_fmm_ex(ct.byref(self.tree._s), ct.byref(self.trav._s),
                                  ct.byref(kp), _ptr(src), _ptr(nrm), _ptr(st),
                                  ctr.shape[1], _ptr(ctr), _ptr(rad), _ptr(tgt), _ptr(ttc),
                                  _ptr(out), _ptr(times), _ptr(mp), _ptr(lc), _ptr(qx))
        assert rc == 0
        return out[self.ncenters:], {"mpole": mp, "local": lc, "qbx": qx, "times": times}

    def direct(self, kernel, k, source_kind, qbx_order, strengths, target_idx=None,
               nthreads=0, target_dir=None):
        tg = self.user_targets
        ttc = self.target_to_center
        td = None if target_dir is None else np.asarray(target_dir, np.float64)
        if target_idx is not None:
            tg = tg[:, target_idx]
            ttc = ttc[target_idx]
            if td is not None:
                td = td[:, target_idx]
        kp = kparams(kernel, k, source_kind, qbx_order, [0], nthreads,
                     self.discr.nodes.shape[0], td)
        return direct_qbx(kp, self.discr.quad_nodes, self.discr.quad_normals,
                          strengths, self.centers, self.radii, tg, ttc)


def assoc_status_and_flags(discr, targets, sides,
                           target_association_tolerance=1.0e3 * 2.220446049250313e-16):
    """(status, refine_flags) for ``problems.refine_until_associated``."""
    ttc, status = target_assoc(discr.nodes, discr.tunnel_radii(), discr.centers(),
                               discr.center_radii() * (1 + target_association_tolerance),
                               targets, sides)
    flags = refine_flags_for_failed(discr.nodes, discr.tunnel_radii(),
                                    discr.element_to_source_starts(), targets, status) \
        if np.any(status == 1) else np.zeros(discr.nelements, np.int32)
    return status, flags


def solid_harmonics(p, x):
    """Regular Y_n^m and irregular T_n^m at x (3,), coefficient (n, m) at
    n*n + n + m (fmm3d_oracle.c)."""
    x = _f64(x)
    Y = np.zeros((p + 1) ** 2, np.complex128)
    T = np.zeros((p + 1) ** 2, np.complex128)
    lib().orc_sh3d(p, _ptr(x), _ptr(Y), _ptr(T))
    return Y, T


def sph_jn(nmax, z):
    """Spherical Bessel j_n(z), n = 0..nmax, and derivatives (complex z)."""
    j = np.zeros(nmax + 1, np.complex128)
    jd = np.zeros(nmax + 1, np.complex128)
    lib().orc_sph_jn(nmax, complex(z).real, complex(z).imag, 1.0, _ptr(j), _ptr(jd))
    return j, jd


def sph_hn(nmax, z):
    """Spherical Hankel h_n^(1)(z), n = 0..nmax, and derivatives (complex z)."""
    h = np.zeros(nmax + 1, np.complex128)
    hd = np.zeros(nmax + 1, np.complex128)
    lib().orc_sph_hn(nmax, complex(z).real, complex(z).imag, 1.0, _ptr(h), _ptr(hd))
    return h, hd


TSQBX_OPS = {"S": 0, "D": 1, "Sp": 2}


def tsqbx(kernel, op, k, order, sources, targets, centers, qbx_centers, c2tb,
          ctt_starts, ctt_lists, l1_starts, l1_lists, box_src_starts, box_src_counts,
          charge, dipvec=None):
    """Target-specific QBX (impl.pyx:566-785) on tree-ordered arrays.
    Returns pot [nt] (S, D) or grad [3, nt] (Sp), complex."""
    src, tgt, ctr = _f64(sources), _f64(targets), _f64(centers)
    ints = [_i32(a) for a in (qbx_centers, c2tb, ctt_starts, ctt_lists, l1_starts,
                              l1_lists, box_src_starts, box_src_counts)]
    q = _c128(charge)
    dv = _f64(dipvec) if dipvec is not None else None
    nt = tgt.shape[1]
    pot = np.zeros(nt, np.complex128)
    grad = np.zeros((3, nt), np.complex128)
    rc = lib().orc_tsqbx(0 if kernel == "laplace" else 1, TSQBX_OPS[op], complex(k).real,
                         complex(k).imag, order, src.shape[1], _ptr(src), nt, _ptr(tgt),
                         ctr.shape[1], _ptr(ctr), len(ints[0]), *[_ptr(a) for a in ints],
                         _ptr(q), _ptr(dv), _ptr(pot), _ptr(grad))
    assert rc == 0
    return grad if op == "Sp" else pot


# ---- 3D Helmholtz point sums at smooth targets (the differentiable path)

def helmholtz3d_dense(sources, targets, k):
    """G_k[t, s] = e^{ik|t-s|} / (4 pi |t-s|) for SoA (3, n) point sets: the
    matrix of sumpy's HelmholtzKernel(3) between quad-stage-2 nodes and
    smooth targets (pytential/symbolic/matrix.py:376-522 builds its entries
    from the same kernel).  Coincident points give zero."""
    s = np.asarray(sources, np.float64)
    t = np.asarray(targets, np.float64)
    d = t[:, :, None] - s[:, None, :]
    r = np.sqrt(np.sum(d * d, axis=0))
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(r > 0, np.exp(1j * k * r) / (4 * np.pi * np.where(r > 0, r, 1.0)), 0.0)
    return g


def helmholtz3d_far_operator(discr, targets, k):
    """A = G_k W U (nodal density -> potential at smooth targets), dense."""
    G = helmholtz3d_dense(discr.quad_nodes, targets, k)
    U = np.kron(np.eye(discr.nelements), discr.upsample)
    return (G * discr.quad_weights[None, :]) @ U


def _upsample_blocks(discr, v):
    """U v for nodal v [ndofs] -> quad-stage-2 values, element by element."""
    ne, n1 = discr.nelements, discr.nodes_per_element
    return (np.asarray(v).reshape(ne, n1) @ discr.upsample.T).reshape(-1)


def helmholtz3d_far_apply(discr, targets, k, sigma, chunk=256):
    """A sigma with A = G_k W U, the rows of G_k built ``chunk`` targets at a
    time (the dense operator of ``helmholtz3d_far_operator`` without storing
    it)."""
    q = discr.quad_weights * _upsample_blocks(discr, sigma)
    t = np.asarray(targets, np.float64)
    out = np.empty(t.shape[1], np.complex128)
    for a in range(0, t.shape[1], chunk):
        out[a:a + chunk] = helmholtz3d_dense(discr.quad_nodes, t[:, a:a + chunk], k) @ q
    return out


def helmholtz3d_far_adjoint(discr, targets, k, y, chunk=256):
    """A^H y = U^T W conj(G_k)^T y, with the rows of G_k built ``chunk``
    targets at a time."""
    t = np.asarray(targets, np.float64)
    acc = np.zeros(discr.quad_nodes.shape[1], np.complex128)
    for a in range(0, t.shape[1], chunk):
        acc += helmholtz3d_dense(discr.quad_nodes, t[:, a:a + chunk], k).conj().T @ y[a:a + chunk]
    acc *= discr.quad_weights
    ne, n1 = discr.nelements, discr.nodes_per_element
    return (acc.reshape(ne, -1) @ discr.upsample).reshape(ne * n1)


def sphere_single_layer_exterior(l, k, a, r):
    """Exterior value of the Helmholtz single layer of Y_lm on the sphere of
    radius a at distance r > a: S_k Y_lm = i k a^2 j_l(k a) h_l(k r) Y_lm,
    from the addition theorem e^{ik|x-y|}/(4 pi |x-y|) = i k sum_l j_l(k a)
    h_l(k r) sum_m Y_lm(x^) conj(Y_lm(y^)) (|y| = a < |x| = r)."""
    from scipy.special import spherical_jn, spherical_yn
    h = spherical_jn(l, k * r) + 1j * spherical_yn(l, k * r)
    return 1j * k * a * a * spherical_jn(l, k * a) * h
