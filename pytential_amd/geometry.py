"""GPU tree, traversal and QBX geometry data (host mirror of boxtree's Tree /
FMMTraversalInfo and of pytential's QBXFMMGeometryData).

Every array lives in HBM; the objects here hold opaque library handles and
export boxtree-named fields as torch tensors on demand.
"""
from __future__ import annotations

import ctypes as ct
from functools import cached_property

import numpy as np
import torch

from . import _lib
from ._lib import call, lib


def _p(t):
    return ct.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    """A library context: owns the FMM workspaces, translation tables and
    scratch of the calls made through it.  ``Context.get`` returns the one
    shared per device; an operator captured into a HIP graph gets its own
    (``Context(device)``), because a graph replays the workspace pointers and
    table contents it captured and the library refuses to replace them once
    captured."""
    _by_device: dict = {}

    def __init__(self, device=0):
        self.device = device
        h = ct.c_void_p()
        call("qbx_ctx_create", ct.byref(h), device)
        self.h = h

    @classmethod
    def get(cls, device=None):
        if device is None:
            device = torch.cuda.current_device()
        if device not in cls._by_device:
            cls._by_device[device] = Context(device)
        return cls._by_device[device]

    def release_graphs(self):
        """No graph captured through this context is replayed any more: its
        workspaces and tables may be replaced again."""
        call("qbx_ctx_release_graphs", self.h)

    def __del__(self):
        try:
            lib().qbx_ctx_destroy(self.h)
        except Exception:
            pass


def as_dev(a, dtype=torch.float64):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


class _Exported:
    """Lazy export of named device fields into torch tensors."""
    _destroy = ""
    _export = ""
    _bytes = ""

    def field(self, name, dtype=torch.int32):
        cache = self.__dict__.setdefault("_fields", {})
        if name not in cache:
            nbytes = getattr(lib(), self._bytes)(self.h, name.encode())
            if nbytes < 0:
                raise KeyError(name)
            esize = torch.tensor([], dtype=dtype).element_size()
            out = torch.empty(max(nbytes // esize, 0), dtype=dtype, device="cuda")
            if nbytes:
                call(self._export, self.h, name.encode(), _p(out), nbytes, _stream())
            cache[name] = out
        return cache[name]

    def __del__(self):
        try:
            getattr(lib(), self._destroy)(self.h)
        except Exception:
            pass


class Tree(_Exported):
    """Adaptive tree built on the GPU (boxtree ``Tree`` field names)."""
    _destroy, _export, _bytes = "qbx_tree_destroy", "qbx_tree_export", "qbx_tree_field_bytes"

    def __init__(self, sources, targets, target_radii=None, source_weights=None,
                 target_weights=None, max_leaf_refine_weight=64, stick_out_factor=0.5,
                 extent_norm="l2", ctx=None):
        self.ctx = ctx or Context.get()
        src = as_dev(sources)
        tgt = as_dev(targets)
        self.dimensions = src.shape[0]
        self._keep = [src, tgt]
        tr = None if target_radii is None else as_dev(target_radii)
        sw = None if source_weights is None else as_dev(source_weights, torch.int32)
        tw = None if target_weights is None else as_dev(target_weights, torch.int32)
        h = ct.c_void_p()
        call("qbx_tree_build", self.ctx.h, _stream(), self.dimensions, src.shape[1], _p(src),
             tgt.shape[1], _p(tgt), _p(tr), _p(sw), _p(tw), max_leaf_refine_weight,
             float(stick_out_factor), int(extent_norm == "linf"), ct.byref(h))
        self.h = h
        sizes = (ct.c_int * 5)()
        root = (ct.c_double * 4)()
        call("qbx_tree_sizes", h, sizes, root)
        _, self.nsources, self.ntargets, self.nboxes, self.nlevels = list(sizes)
        self.bounding_box_min = np.array(root[:self.dimensions])
        self.root_extent = root[3]

    def __getattr__(self, name):
        if name.startswith("_") or name in ("h", "ctx"):
            raise AttributeError(name)
        dtype = torch.float64 if name in ("box_centers", "sources", "targets") else torch.int32
        try:
            t = self.field(name, dtype)
        except KeyError:
            raise AttributeError(name) from None
        if name in ("box_centers", "box_coords", "sources", "targets"):
            return t.view(self.dimensions, -1)
        if name == "box_child_ids":
            return t.view(2 ** self.dimensions, -1)
        return t


class Traversal(_Exported):
    """FMM interaction lists built on the GPU (boxtree FMMTraversalInfo names,
    close lists merged into list 1)."""
    _destroy, _export, _bytes = "qbx_trav_destroy", "qbx_trav_export", "qbx_trav_field_bytes"

    def __init__(self, tree: Tree, well_sep_is_n_away=2, targets_have_extent=True,
                 from_sep_smaller_min_nsources_cumul=15):
        self.tree = tree
        h = ct.c_void_p()
        call("qbx_traverse", tree.ctx.h, _stream(), tree.h, well_sep_is_n_away,
             int(targets_have_extent), from_sep_smaller_min_nsources_cumul, ct.byref(h))
        self.h = h
        c = (ct.c_int * 8)()
        call("qbx_trav_sizes", h, c)
        (self.nsource_boxes, self.nsource_parent_boxes, self.ntarget_boxes,
         self.ntarget_or_target_parent_boxes) = c[:4]

    def __getattr__(self, name):
        if name.startswith("_") or name in ("h", "tree"):
            raise AttributeError(name)
        try:
            return self.field(name)
        except KeyError:
            raise AttributeError(name) from None


class QBXTargetAssociationFailedError(RuntimeError):
    """Mirror of pytential.qbx.target_assoc.QBXTargetAssociationFailedError."""

    def __init__(self, refine_flags, failed_target_flags, message):
        super().__init__(message)
        self.refine_flags = refine_flags
        self.failed_target_flags = failed_target_flags
        self.message = message


TARGET_ASSOC_TOLERANCE = 1.0e3 * float(np.finfo(np.float64).eps)


def associate_targets_to_qbx_centers(discr_dev, targets, flags,
                                     target_association_tolerance=TARGET_ASSOC_TOLERANCE,
                                     ctx=None):
    """GPU target association (pytential/qbx/target_assoc.py:853-952).
    Returns ``target_to_center`` (int32, -1 = no QBX needed)."""
    ctx = ctx or Context.get()
    nodes, tunnel, centers, radii, el_starts = (
        discr_dev["nodes"], discr_dev["tunnel"], discr_dev["centers"],
        discr_dev["center_radii"], discr_dev["element_to_source_starts"])
    tg = as_dev(targets)
    fl = as_dev(flags, torch.int32)
    nt = tg.shape[1]
    rtol = radii * (1 + target_association_tolerance)
    ttc = torch.empty(nt, dtype=torch.int32, device="cuda")
    status = torch.empty(nt, dtype=torch.int32, device="cuda")
    nfailed = ct.c_int(0)
    rc = lib().qbx_target_assoc(ctx.h, _stream(), nodes.shape[0], nodes.shape[1], _p(nodes),
                                _p(tunnel), centers.shape[1], _p(centers), _p(rtol), nt,
                                _p(tg), _p(fl), _p(ttc), _p(status), ct.byref(nfailed))
    if rc == _lib.QBX_ERR_TARGET_ASSOC_FAILED:
        nel = el_starts.shape[0] - 1
        refine = torch.empty(nel, dtype=torch.int32, device="cuda")
        call("qbx_refine_flags_for_failed", ctx.h, _stream(), nodes.shape[0], nodes.shape[1],
             _p(nodes), _p(tunnel), nel, _p(el_starts), nt, _p(tg), _p(status), _p(refine))
        surface = (fl.abs() == 1) & (status == 1)
        msg = ("An on-surface target was not assigned a QBX center."
               if bool(surface.any()) else "Some targets were not assigned a QBX center.")
        raise QBXTargetAssociationFailedError(refine, status == 1, msg)
    _lib.check(rc)
    return ttc, status


class QBXFMMGeometryData:
    """Mirror of pytential.qbx.geometry.QBXFMMGeometryData for one source
    curve or surface and a list of ``(targets, qbx_side)`` pairs; everything on the GPU."""

    def __init__(self, lpot_source, target_discrs_and_qbx_sides):
        self.lpot_source = lpot_source
        self.target_discrs_and_qbx_sides = target_discrs_and_qbx_sides
        self.ctx = lpot_source.ctx
        self.dev = lpot_source.dev

    @property
    def ncenters(self):
        return self.dev["centers"].shape[1]

    def flat_centers(self):
        return self.dev["centers"]

    def flat_expansion_radii(self):
        return self.dev["center_radii"]

    @cached_property
    def _target_info(self):
        pts, sides, starts = [], [], []
        n = self.ncenters
        for tgt, side in self.target_discrs_and_qbx_sides:
            t = self.dev["nodes"] if isinstance(tgt, str) and tgt == "self" else as_dev(tgt)
            starts.append(n)
            n += t.shape[1]
            pts.append(t)
            s = torch.as_tensor(side, dtype=torch.int32, device="cuda")
            sides.append(s.expand(t.shape[1]) if s.dim() == 0 else s)
        starts.append(n)
        user = torch.cat(pts, dim=1) if pts else torch.zeros(
            (self.lpot_source.ambient_dim, 0), dtype=torch.float64, device="cuda")
        side = torch.cat(sides) if sides else torch.zeros(0, dtype=torch.int32, device="cuda")
        return {"user_targets": user.contiguous(), "sides": side.contiguous(),
                "target_discr_starts": starts}

    def target_info(self):
        ti = self._target_info
        return {"targets": torch.cat([self.flat_centers(), ti["user_targets"]], dim=1),
                "target_discr_starts": ti["target_discr_starts"],
                "ntargets": ti["target_discr_starts"][-1]}

    @cached_property
    def _assoc(self):
        ti = self._target_info
        return associate_targets_to_qbx_centers(
            self.dev, ti["user_targets"], ti["sides"],
            self.lpot_source.target_association_tolerance, self.ctx)

    def user_target_to_center(self):
        ttc, _ = self._assoc
        return torch.cat([torch.full((self.ncenters,), -1, dtype=torch.int32, device="cuda"),
                          ttc])

    @cached_property
    def _tree(self):
        ls = self.lpot_source
        ti = self.target_info()
        uttc = self.user_target_to_center()
        radii = torch.zeros(ti["ntargets"], dtype=torch.float64, device="cuda")
        radii[:self.ncenters] = self.flat_expansion_radii()
        tw = torch.ones(ti["ntargets"], dtype=torch.int32, device="cuda")
        tw[self.ncenters:] = (uttc[self.ncenters:] == -1).to(torch.int32)
        return Tree(self.dev["quad_nodes"], ti["targets"], radii, None, tw,
                    ls._max_leaf_refine_weight, ls._expansion_stick_out_factor,
                    ls._box_extent_norm, ctx=self.ctx)

    def tree(self):
        return self._tree

    @cached_property
    def _traversal(self):
        ls = self.lpot_source
        return Traversal(self._tree, ls._well_sep_is_n_away, True,
                         ls._from_sep_smaller_min_nsources_cumul)

    def traversal(self):
        return self._traversal

    @cached_property
    def _geo(self):
        uttc = self.user_target_to_center().contiguous()
        h = ct.c_void_p()
        call("qbx_geo_build", self.ctx.h, _stream(), self._tree.h, self._traversal.h,
             self.ncenters, _p(self.flat_centers()), _p(self.flat_expansion_radii()),
             _p(uttc), ct.byref(h))
        g = _GeoHandle(h)
        g._keep = uttc
        return g

    def qbx_center_to_target_box(self):
        return self._geo.field("qbx_center_to_target_box")

    def global_qbx_centers(self):
        return self._geo.field("global_qbx_centers")

    def center_to_tree_targets(self):
        return (self._geo.field("center_to_tree_targets_starts"),
                self._geo.field("center_to_tree_targets_lists"))

    def non_qbx_box_target_lists(self):
        g = self._geo
        return {"box_target_starts": g.field("nqbtl_box_target_starts"),
                "box_target_counts_nonchild": g.field("nqbtl_box_target_counts_nonchild"),
                "unfiltered_from_filtered_target_indices":
                    g.field("nqbtl_unfiltered_from_filtered_target_indices")}

    def build_all(self):
        """Force tree, traversal, association and lookups (the 'cold' part)."""
        return self._geo


class _GeoHandle(_Exported):
    _destroy, _export, _bytes = "qbx_geo_destroy", "qbx_geo_export", "qbx_geo_field_bytes"

    def __init__(self, h):
        self.h = h
        c = (ct.c_int * 4)()
        call("qbx_geo_sizes", h, c)
        self.ncenters, self.nglobal, self.nctt, self.nfiltered = c[:]


def assoc_status_and_flags(discr, targets, sides,
                           target_association_tolerance=TARGET_ASSOC_TOLERANCE):
    """GPU (status, refine_flags) of a discretization against targets, for
    ``problems.refine_until_associated``."""
    dev = {"nodes": as_dev(discr.nodes), "tunnel": as_dev(discr.tunnel_radii()),
           "centers": as_dev(discr.centers()), "center_radii": as_dev(discr.center_radii()),
           "element_to_source_starts": as_dev(discr.element_to_source_starts(), torch.int32)}
    try:
        _, status = associate_targets_to_qbx_centers(dev, targets, sides,
                                                     target_association_tolerance)
        return status.cpu().numpy(), np.zeros(discr.nelements, np.int32)
    except QBXTargetAssociationFailedError as e:
        st = e.failed_target_flags.cpu().numpy().astype(np.int32)
        return st, e.refine_flags.cpu().numpy()
