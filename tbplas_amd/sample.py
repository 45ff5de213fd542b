"""Drop-in ``Sample`` for the TBPM path, and sample generators.

``Sample`` holds the same arrays as the reference's Sample
(tbplas/builder/sample.py:24-106): ``orb_eng``, ``orb_pos``, ``hop_i``,
``hop_j``, ``hop_v``, ``dr`` and ``_rescale``.  Its Hamiltonian builders run
on the GPU through the C ABI:

* ``build_ham_csr``  -> scipy-canonical CSR (sample.py:375-391)
* ``build_ham_dxy``  -> insertion-order CSR + dx, dy (sample.py:417-461)
* ``rescale_ham``    -> sample.py:272-299 (estimate on the GPU when factor
  is None, sample.pyx:75-114)
* ``set_magnetic_field`` -> Peierls phases (sample.py:301-354,
  sample.pyx:18-72), in the library's host code with libm cos/sin.

A Sample is built from arrays, from a reference Sample (``from_tbplas``), from
files written by the reference's ``save_array`` (``load_array``), or by the
graphene generator ``make_graphene_sample`` below.
"""
import math
import os

import numpy as np

from . import _lib

__all__ = ["Sample", "make_graphene_sample", "graphene_hops", "supercell_hops",
           "add_anderson_disorder", "apply_uniaxial_strain", "random_vacancies",
           "auto_tile_nm", "ac_tile_nm"]

GRAPHENE_A = 0.246      # nm, make_graphene_diamond (materials/graphene.py:16-31)
GRAPHENE_T = -2.7       # eV


def graphene_hops(L1, L2, t=GRAPHENE_T, pbc=(True, True), vacancies=None,
                  a=GRAPHENE_A):
    """Graphene supercell arrays in the reference's order (library host
    code, ``tbg_graphene_hops``).

    Primitive cell of make_graphene_diamond (materials/graphene.py:16-31):
    lattice a at 60 deg, orbitals at (0,0) and (1/3,1/3), reduced hops
    [(0,0)0->1, (1,0)1->0, (0,1)1->0] (order from check_conj,
    builder/base.py:76-92).  Orbital order (ia, ib, io) (super.pyx:383-443,
    vacancies skipped).  Hop order: per cell (ia, ib), the A->B hop then the
    two B->A hops, as build_hop_pbc / build_hop_gen emit them for this cell
    (super.pyx:643-986); hops touching a vacancy are dropped; for open
    directions the crossing hops come after all periodic ones
    (split_pc_hop, super.pyx:786-857).
    Returns orb_pos (n,3), hop_i, hop_j, hop_v, dr (m,3).
    """
    import ctypes
    L1, L2 = int(L1), int(L2)
    n0, m0 = 2 * L1 * L2, 3 * L1 * L2
    vac = (np.zeros(0, dtype=np.int64) if vacancies is None
           else np.ascontiguousarray(vacancies, dtype=np.int64))
    pos = np.empty((n0, 3))
    hi = np.empty(m0, dtype=np.int64)
    hj = np.empty(m0, dtype=np.int64)
    hv = np.empty(m0, dtype=np.complex128)
    dr = np.empty((m0, 3))
    n, m = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.lib().tbg_graphene_hops(
        L1, L2, float(t), int(bool(pbc[0])), int(bool(pbc[1])), float(a),
        vac.ctypes.data if len(vac) else None, len(vac), pos.ctypes.data, hi.ctypes.data,
        hj.ctypes.data, hv.ctypes.data, dr.ctypes.data, ctypes.byref(n), ctypes.byref(m)))
    n, m = n.value, m.value
    return pos[:n].copy(), hi[:m].copy(), hj[:m].copy(), hv[:m].copy(), dr[:m].copy()


GRAPHENE_ORB_DENSITY = 38.2   # orbitals per nm^2 (2 per 0.0524 nm^2 cell)


def pass_terms(vectors, precision="fp32"):
    """Terms per pass the tiles of ``auto_tile_nm`` are sized for: five
    where a row of the device layout takes at most 128 B (complex64 up to
    16 states, complex128 up to 8), else four.  Five terms hold two more
    rings in LDS, which shrinks the tile; with rows this short the tile
    stays large enough that the halo costs less than the pass saves.
    Measured on 8.4M-orbital graphene (step·vectors/s, five terms against
    four): complex64 K=8 1904 at 4.8 nm against 1724 at 5.0 nm; K=4 1803
    at 7.1 nm against 1620 at 7.4 nm; K=16 1822 at 3.1 nm against 1777 at
    3.4 nm; complex128 K=1 658 against 588 at 7.5 nm; K=4 915 at 4.8 nm
    against 805 at 5.0 nm; K=8 896 at 3.1 nm against 848 at 3.4 nm."""
    return 5 if _row_bytes(vectors, precision) <= 128 else 4


def _row_bytes(vectors, precision):
    p2, p3 = 1, 3            # vector slots of the device layout
    while p2 < vectors:
        p2 <<= 1
    while p3 < vectors:
        p3 <<= 1
    return min(p2, p3) * (8 if precision == "fp32" else 16)


def auto_tile_nm(vectors, precision="fp32", density=GRAPHENE_ORB_DENSITY, terms=None,
                 n_orb=None, cus=None):
    """Edge (nm) of the square tiles of ``Sample.tile_order`` for runs of
    ``vectors`` states per row on a sample of ``density`` orbitals per nm^2
    with ``terms`` Chebyshev terms per pass (default ``pass_terms``): the
    largest tile whose rows fit 150 KB of LDS, within [1.5, 7.5] nm.  The
    LDS areas A and B hold the tile twice and its rings 2 T - 3 times (four
    terms: the first ring twice, the second twice, the third once; five
    terms two rings more), with 2 rho a^2 tile rows and a ring of
    2.96 sqrt(rho) a rows (nearest-neighbour lattices; fitted on graphene,
    17.6 a + 2 per ring).  A tile any larger is split by the library into
    strips, which costs more than it saves.  Measured with four terms on
    8.4M-orbital graphene (step·vectors/s): complex64 K=8 1607 at 5.0 nm,
    1458 at 3.5 nm; K=16 1643 at 3.0 nm, 1436 at 3.54 nm (split);
    complex128 K=1 (capped) 591 steps/s at 7.0 nm, 587 at 5.0, 570 at 9.0;
    and a batch of 32 disordered samples with complex64 K=4 (7.35 nm)
    13426 instance-step·vectors/s at 7.0 nm against 11720 at 5.0.

    Given the sample's ``n_orb`` orbitals and the GPU's ``cus`` compute
    units, a small sample is cut into at least two tiles per CU (never
    below 1.5 nm), so that a pass has enough workgroups: on 131,072-orbital
    graphene with one state (2.59 nm) complex128 runs 41.6 us per step on
    2.5 nm tiles, 49.8 on 2.0 and 45.8 on 3.0, against 68 with one term per
    pass, which the tuner takes on 7.5 nm tiles (61 of them)."""
    if terms is None:
        terms = pass_terms(vectors, precision)
    rows = 150.0 * 1024 / _row_bytes(vectors, precision)   # LDS rows of one workgroup
    a2 = 2.0 * density                            # tile rows per nm^2, A and B
    a1 = (2 * terms - 3) * 2.96 * math.sqrt(density)   # ring rows per nm of edge
    edge = (-a1 + math.sqrt(a1 * a1 + 4.0 * a2 * rows)) / (2.0 * a2)
    if n_orb is not None and cus:
        edge = min(edge, math.sqrt(n_orb / density / (2.0 * cus)))
    return float(min(7.5, max(1.5, edge)))


def ac_tile_nm(vectors, precision="fp32", density=GRAPHENE_ORB_DENSITY, n_orb=None, cus=None):
    """Tile edge (nm) of AC runs with ``vectors`` propagated vectors per row
    (three per state): complex64 takes ``auto_tile_nm``; complex128, whose
    three states per lane run 8-wave workgroups that gain from two per CU,
    takes the smaller of that and 2.5 nm at graphene's density (scaled to
    the same rows per tile at others).  Measured on 8.4M-orbital graphene,
    4 states a run (state·steps/s): complex64 753 on 3.8 nm (the rule),
    725 on 3.6, 645 on 3.0, 648 on 2.5; complex128 376 on 2.5, 336 on 2.7,
    335 on 3.0, 332 on 3.8."""
    edge = auto_tile_nm(vectors, precision, density, n_orb=n_orb, cus=cus)
    if precision == "fp64":
        edge = min(edge, max(1.5, 2.5 * math.sqrt(GRAPHENE_ORB_DENSITY / density)))
    return float(edge)


def random_vacancies(n_orb, fraction, seed):
    """Sorted orbital indices of a seeded uniform vacancy draw."""
    rng = np.random.default_rng(seed)
    nv = int(round(fraction * n_orb))
    return np.sort(rng.choice(n_orb, nv, replace=False)).astype(np.int64)


class Sample:
    """Arrays of a tight-binding sample plus GPU Hamiltonian builders."""

    def __init__(self, orb_eng=None, orb_pos=None, hop_i=None, hop_j=None,
                 hop_v=None, dr=None, rescale=1.0, area_unit_cell=None,
                 volume_unit_cell=None, extended=1, nr_orbitals=None,
                 device=0):
        self.orb_eng = None if orb_eng is None else np.ascontiguousarray(orb_eng, dtype=np.float64)
        self.orb_pos = None if orb_pos is None else np.ascontiguousarray(orb_pos, dtype=np.float64)
        self.hop_i = None if hop_i is None else np.ascontiguousarray(hop_i, dtype=np.int64)
        self.hop_j = None if hop_j is None else np.ascontiguousarray(hop_j, dtype=np.int64)
        self.hop_v = None if hop_v is None else np.ascontiguousarray(hop_v, dtype=np.complex128)
        self.dr = None if dr is None else np.ascontiguousarray(dr, dtype=np.float64).reshape(-1, 3)
        self._rescale = float(rescale)
        self._area_unit_cell = area_unit_cell
        self._volume_unit_cell = volume_unit_cell
        self._extended = extended
        self._nr_orbitals = nr_orbitals
        self.device = device

    # -- construction -----------------------------------------------------
    @classmethod
    def from_tbplas(cls, sample, device=0):
        """Copy the arrays of a reference ``tbplas.Sample``."""
        sample.init_orb_eng()
        sample.init_orb_pos()
        sample.init_hop()
        return cls(sample.orb_eng, sample.orb_pos, sample.hop_i, sample.hop_j,
                   sample.hop_v, sample.dr, sample.rescale,
                   area_unit_cell=sample.area_unit_cell,
                   volume_unit_cell=sample.volume_unit_cell,
                   extended=sample.extended, nr_orbitals=sample.nr_orbitals,
                   device=device)

    # -- reference API ----------------------------------------------------
    def init_orb_eng(self, force_init=False):
        if self.orb_eng is None:
            raise ValueError("orb_eng not set")

    def init_orb_pos(self, force_init=False):
        if self.orb_pos is None:
            raise ValueError("orb_pos not set")

    def init_hop(self, force_init=False):
        if self.hop_i is None:
            raise ValueError("hopping arrays not set")

    def tile_order(self, tile=4.0, return_tiles=False):
        """Row order for the device layout (``Hamiltonian(order=...)``):
        orbitals binned into square tiles of ``tile`` nm in the x-y plane,
        tiles row by row, the original order inside a tile.  Neighbours then
        mostly share a 64-row slice or sit in the next tile, instead of a
        whole lattice row away.  With ``return_tiles`` also the row offsets
        of the non-empty tiles in that order (``Hamiltonian(tiles=...)``)."""
        self.init_orb_pos()
        xy = self.orb_pos[:, :2]
        lo = xy.min(axis=0)
        b = np.floor((xy - lo) / float(tile)).astype(np.int64)
        nx = int(b[:, 0].max()) + 1
        key = b[:, 1] * nx + b[:, 0]
        perm = np.argsort(key, kind="stable").astype(np.int64)
        if not return_tiles:
            return perm
        ks = key[perm]
        starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
        return perm, np.r_[starts, len(perm)].astype(np.int64)

    def save_array(self, data_dir="sample"):
        """Same files as the reference (sample.py:220-240)."""
        os.makedirs(data_dir, exist_ok=True)
        if self.orb_eng is not None:
            np.save(f"{data_dir}/orb_eng", self.orb_eng)
        if self.orb_pos is not None:
            np.save(f"{data_dir}/orb_pos", self.orb_pos)
        if self.hop_i is not None:
            np.save(f"{data_dir}/hop_i", self.hop_i)
            np.save(f"{data_dir}/hop_j", self.hop_j)
            np.save(f"{data_dir}/hop_v", self.hop_v)
            np.save(f"{data_dir}/dr", self.dr)
            np.save(f"{data_dir}/rescale", self._rescale)

    def load_array(self, data_dir="sample"):
        """Read files written by save_array (sample.py:242-262); numpy
        loads with allow_pickle=False."""
        def _ld(name):
            return np.load(f"{data_dir}/{name}.npy", allow_pickle=False)
        for name in ("orb_eng", "orb_pos"):
            try:
                setattr(self, name, _ld(name))
            except FileNotFoundError:
                print(f"Ignoring {data_dir}/{name}.npy")
        try:
            self.hop_i, self.hop_j = _ld("hop_i"), _ld("hop_j")
            self.hop_v, self.dr = _ld("hop_v"), _ld("dr")
            self._rescale = float(_ld("rescale").item())
        except FileNotFoundError:
            print(f"Ignoring {data_dir}/hop_i.npy")

    def rescale_ham(self, factor=None):
        """sample.py:272-299: divide by factor relative to the originals."""
        self.init_orb_eng()
        self.init_hop()
        if factor is None:
            factor = self.get_rescale()
        self.orb_eng *= (self._rescale / factor)
        self.hop_v *= (self._rescale / factor)
        self._rescale = factor

    def get_rescale(self):
        """Max row sum of |H| (sample.pyx:75-114) on the GPU."""
        import ctypes
        h = _lib.default_handle(self.device)
        d = _DevHops(self, self.device)
        out = ctypes.c_double(0.0)
        _lib.check(_lib.lib().tbg_get_rescale(
            h.ptr, self.num_orb, self.num_hop, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
            _lib.ptr(d.hop_j), _lib.ptr(d.hop_v), ctypes.byref(out)))
        return out.value

    def set_magnetic_field(self, intensity, gauge=0):
        """Peierls substitution (sample.py:301-354, sample.pyx:18-72)."""
        self.init_orb_pos()
        self.init_hop()
        if gauge not in (0, 1, 2):
            raise ValueError(f"Illegal gauge {gauge}")
        _lib.check(_lib.lib().tbg_peierls_host(
            self.num_hop, self.hop_i.ctypes.data, self.hop_j.ctypes.data,
            self.hop_v.ctypes.data, self.dr.ctypes.data, self.orb_pos.ctypes.data,
            float(intensity), int(gauge)))

    def build_ham_csr_device(self):
        """Canonical CSR on the GPU: (indptr, indices, data) device tensors."""
        import ctypes
        torch = _lib.torch_cuda()
        self.init_orb_eng()
        self.init_hop()
        h = _lib.default_handle(self.device)
        d = _DevHops(self, self.device)
        n, m = self.num_orb, self.num_hop
        dev = f"cuda:{self.device}"
        indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        nnz = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.tbg_ham_csr_count(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                       _lib.ptr(d.hop_j), _lib.ptr(d.hop_v), _lib.ptr(indptr),
                                       ctypes.byref(nnz)))
        indices = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=dev)
        data = torch.empty(max(nnz.value, 1), dtype=torch.complex128, device=dev)
        _lib.check(L.tbg_ham_csr_fill(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                      _lib.ptr(d.hop_j), _lib.ptr(d.hop_v), _lib.ptr(indptr),
                                      _lib.ptr(indices), _lib.ptr(data)))
        return indptr, indices[:nnz.value], data[:nnz.value]

    def build_ham_csr(self):
        """Canonical sparse Hamiltonian as scipy.sparse.csr_matrix."""
        from scipy.sparse import csr_matrix
        ip, ix, dt = self.build_ham_csr_device()
        n = self.num_orb
        return csr_matrix((dt.cpu().numpy(), ix.cpu().numpy().astype(np.int32),
                           ip.cpu().numpy().astype(np.int32)), shape=(n, n))

    def build_ham_dxy_device(self, orb_eng_cutoff=1.0e-5):
        """Insertion-order CSR + dx, dy on the GPU, as device tensors."""
        import ctypes
        torch = _lib.torch_cuda()
        self.init_orb_eng()
        self.init_orb_pos()
        self.init_hop()
        h = _lib.default_handle(self.device)
        d = _DevHops(self, self.device, with_dr=True)
        n, m = self.num_orb, self.num_hop
        dev = f"cuda:{self.device}"
        indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        nnz = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.tbg_ham_dxy_count(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                       _lib.ptr(d.hop_j), float(orb_eng_cutoff),
                                       _lib.ptr(indptr), ctypes.byref(nnz)))
        k = max(nnz.value, 1)
        indices = torch.empty(k, dtype=torch.int64, device=dev)
        hop = torch.empty(k, dtype=torch.complex128, device=dev)
        dx = torch.empty(k, dtype=torch.float64, device=dev)
        dy = torch.empty(k, dtype=torch.float64, device=dev)
        _lib.check(L.tbg_ham_dxy_fill(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                      _lib.ptr(d.hop_j), _lib.ptr(d.hop_v), _lib.ptr(d.dr),
                                      float(orb_eng_cutoff), _lib.ptr(indptr),
                                      _lib.ptr(indices), _lib.ptr(hop), _lib.ptr(dx),
                                      _lib.ptr(dy)))
        z = nnz.value
        return indptr, indices[:z], hop[:z], dx[:z], dy[:z]

    def build_ham_dxy(self, orb_eng_cutoff=1.0e-5, algo="fast", sort_col=False):
        """(indptr, indices, hop, dx, dy) numpy arrays (sample.py:417-461).

        ``algo`` is accepted for compatibility; the GPU builder always
        produces the "fast" (insertion) order.  ``sort_col`` sorts each row
        by column like sort_col_csr (sample.pyx:385-439).
        """
        if algo not in ("fast", "safe"):
            raise ValueError(f"unknown algo {algo}")
        out = [t.cpu().numpy() for t in self.build_ham_dxy_device(orb_eng_cutoff)]
        if sort_col:
            indptr, indices, hop, dx, dy = out
            for r in range(len(indptr) - 1):
                a, b = indptr[r], indptr[r + 1]
                o = np.argsort(indices[a:b], kind="stable")
                for arr in (indices, hop, dx, dy):
                    arr[a:b] = arr[a:b][o]
        return tuple(out)

    # -- properties (sample.py:644-738) --------------------------------------
    @property
    def rescale(self):
        return self._rescale

    @property
    def num_orb(self):
        return int(len(self.orb_eng))

    @property
    def num_hop(self):
        return int(len(self.hop_i))

    @property
    def energy_range(self):
        return 2.0 * self._rescale

    @property
    def area_unit_cell(self):
        return self._area_unit_cell

    @property
    def volume_unit_cell(self):
        return self._volume_unit_cell

    @property
    def extended(self):
        return self._extended

    @property
    def nr_orbitals(self):
        return self._nr_orbitals


class _DevHops:
    """Device copies of a sample's arrays for one builder call."""

    def __init__(self, s, device, with_dr=False):
        self.orb_eng = _lib.dev(s.orb_eng, np.float64, device)
        self.hop_i = _lib.dev(s.hop_i, np.int64, device)
        self.hop_j = _lib.dev(s.hop_j, np.int64, device)
        self.hop_v = _lib.dev(s.hop_v, np.complex128, device)
        self.dr = _lib.dev(s.dr, np.float64, device) if with_dr else None


def supercell_hops(pc_lattice, pc_orb_frac, pc_hop_ind, pc_hop_eng, dim,
                   pbc=(True, True, True), vacancies=None, device=0):
    """Supercell arrays of a general primitive cell in the reference's
    order, generated on the GPU (``tbg_supercell_hops``).

    pc_lattice (3, 3) Cartesian vectors in nm; pc_orb_frac (n_pc, 3)
    fractional orbital positions; pc_hop_ind (m_pc, 5) reduced hopping terms
    (ra, rb, rc, orb_i, orb_j) as the primitive cell stores them after the
    conjugate reduction, pc_hop_eng (m_pc,) complex; dim (3,) repetitions;
    vacancies: (ia, ib, ic, orb) tuples.  Orbital order (ia, ib, ic, orb)
    with vacancies skipped (super.pyx:383-443), positions as build_orb_pos
    (super.pyx:569-612), hop order of SuperCell.get_hop
    (builder/super.py:703-775).  Returns orb_pos, hop_i, hop_j, hop_v, dr."""
    import ctypes
    lat = np.ascontiguousarray(pc_lattice, dtype=np.float64).reshape(3, 3)
    frac = np.ascontiguousarray(pc_orb_frac, dtype=np.float64).reshape(-1, 3)
    ind = np.ascontiguousarray(pc_hop_ind, dtype=np.int32).reshape(-1, 5)
    eng = np.ascontiguousarray(pc_hop_eng, dtype=np.complex128).reshape(-1)
    dim = np.ascontiguousarray(dim, dtype=np.int32).reshape(3)
    pbc = np.ascontiguousarray([int(bool(x)) for x in pbc], dtype=np.int32)
    npc = len(frac)
    cells = np.indices(tuple(int(d) for d in dim)).reshape(3, -1).T      # ia, ib, ic
    full = np.concatenate([np.repeat(cells, npc, axis=0),
                           np.tile(np.arange(npc), len(cells))[:, None]], axis=1)
    if vacancies is not None and len(vacancies):
        v = np.asarray(vacancies, dtype=np.int64).reshape(-1, 4)
        vac = np.unique(((v[:, 0] * dim[1] + v[:, 1]) * dim[2] + v[:, 2]) * npc + v[:, 3])
    else:
        vac = np.zeros(0, dtype=np.int64)
    keep = np.ones(len(full), dtype=bool)
    keep[vac] = False
    orb_id = np.ascontiguousarray(full[keep], dtype=np.int32)
    r = frac[orb_id[:, 3]] + orb_id[:, :3]
    pos = np.empty((len(orb_id), 3))
    for x in range(3):
        pos[:, x] = r[:, 0] * lat[0, x] + r[:, 1] * lat[1, x] + r[:, 2] * lat[2, x]
    sc_lat = np.ascontiguousarray(lat * dim[:, None].astype(np.float64))
    cap = max(1, len(ind) * int(np.prod(dim)))
    hi = np.empty(cap, dtype=np.int64)
    hj = np.empty(cap, dtype=np.int64)
    hv = np.empty(cap, dtype=np.complex128)
    dr = np.empty((cap, 3))
    m = ctypes.c_int64()
    h = _lib.default_handle(device)
    _lib.check(_lib.lib().tbg_supercell_hops(
        h.ptr, ind.ctypes.data if len(ind) else None, eng.ctypes.data if len(eng) else None,
        len(ind), dim.ctypes.data, pbc.ctypes.data, npc, orb_id.ctypes.data, len(orb_id),
        vac.ctypes.data if len(vac) else None, len(vac), sc_lat.ctypes.data,
        np.ascontiguousarray(pos).ctypes.data, cap, hi.ctypes.data, hj.ctypes.data,
        hv.ctypes.data, dr.ctypes.data, ctypes.byref(m)))
    k = m.value
    return pos, hi[:k].copy(), hj[:k].copy(), hv[:k].copy(), dr[:k].copy()


def make_graphene_sample(L1, L2=None, t=GRAPHENE_T, pbc=(True, True),
                         vacancies=None, device=0):
    """Graphene supercell sample (2 orbitals per cell, a = 0.246 nm)."""
    L2 = L1 if L2 is None else L2
    pos, hi, hj, hv, dr = graphene_hops(L1, L2, t, pbc, vacancies)
    a = GRAPHENE_A
    area = a * a * math.sin(math.pi / 3)
    return Sample(np.zeros(len(pos)), pos, hi, hj, hv, dr, 1.0,
                  area_unit_cell=area, volume_unit_cell=area * 1.0,
                  extended=1, nr_orbitals=2, device=device)


def add_anderson_disorder(sample, W, seed):
    """orb_eng += U[-W/2, W/2] (seeded), applied before rescale_ham."""
    rng = np.random.default_rng(seed)
    sample.orb_eng = sample.orb_eng + rng.uniform(-0.5 * W, 0.5 * W, sample.num_orb)


def apply_uniaxial_strain(sample, eps, axis=0, beta=3.37, d0=None):
    """Stretch coordinates along ``axis`` by (1 + eps) and scale hoppings by
    exp(-beta (d - d0) / d0), d the strained bond length."""
    if d0 is None:
        d0 = GRAPHENE_A / math.sqrt(3.0)
    sample.orb_pos = sample.orb_pos.copy()
    sample.orb_pos[:, axis] *= (1.0 + eps)
    sample.dr = sample.dr.copy()
    sample.dr[:, axis] *= (1.0 + eps)
    d = np.linalg.norm(sample.dr, axis=1)
    sample.hop_v = sample.hop_v * np.exp(-beta * (d - d0) / d0)
