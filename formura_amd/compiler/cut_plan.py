"""Region plan of the reference's MPI C++ backend, restated from
src/Formura/MPICxx/Cut.hs:142-616.  The compiler takes the state arrays'
extents and the navigator's neighbour-rank fields from it (abi.py, as
MPICxx/Translate.hs:935-969 and 1032-1035 do), and the region-plan CPU
library of the oracle (oracle/emit_planned.py) its loop and region structure.
The GPU kernels do not use it: on one GPU the periodic wrap replaces the
ridges and facets.  The grid of each rank
is cut by walls (`initial_walls`) into regions ("IRanks"); each manifest node
of the (temporally blocked) step graph gets its own walls, moved by the
offsets it reads at, so a region can be computed from the regions before it;
what a region reads outside its own rank or region arrives through ridges
(halo slabs) and, across ranks, facets (MPI messages).  The plan lists the
region boxes, the array allocations, the ridges and facets, and the
distributed program (stage / unstage / compute / free / send-receive / wait)
in the reference's order.  Two plans are made: the normal one and the
wall-inverted one (MPICxx/Translate.hs:990-998).

Orders follow the reference's Haskell instances, because they decide names,
array shapes and the instruction order:
  * `Levitated`: Top < Levitate n < Bottom (derived Ord), and an orthotope's
    `&&&` takes max of lower vertices, min of upper vertices
    (Geometry.hs:82-86);
  * `Vec`: lexicographic (Vec.hs:56-57);
  * `IRank`: compares the reversed index vector unless the option
    `irank-order-f` is given (Cut.hs:51-57, 281-283);
  * `ResourceT`: ResourceStatic before ResourceOMNode, then by name / id.
Nanoblock unrolling (`nbu<axis>N` > 1) is not restated.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

# ---- Levitated Int: (0,) Top < (1, n) Levitate n < (2,) Bottom -------------
TOP = (0, 0)
BOT = (2, 0)


def lev(n: int) -> tuple:
    return (1, n)


def lev_move(x: tuple, d: int) -> tuple:
    return (1, x[1] + d) if x[0] == 1 else x


Vec = Tuple[int, ...]


@dataclass(frozen=True)
class Box:
    lo: Vec
    hi: Vec

    def move(self, v: Vec) -> "Box":
        return Box(tuple(a + b for a, b in zip(self.lo, v)), tuple(a + b for a, b in zip(self.hi, v)))

    def meet(self, o: "Box") -> "Box":           # &&&
        return Box(tuple(map(max, self.lo, o.lo)), tuple(map(min, self.hi, o.hi)))

    def hull(self, o: "Box") -> "Box":           # |||
        return Box(tuple(map(min, self.lo, o.lo)), tuple(map(max, self.hi, o.hi)))

    def volume(self) -> int:
        v = 1
        for l, u in zip(self.lo, self.hi):
            v *= (u - l) if u > l else 0
        return v

    def shape(self) -> Vec:
        return tuple(u - l for l, u in zip(self.lo, self.hi))

    def key(self):
        return (self.lo, self.hi)


# A partition (wall) is an orthotope over Levitated Int: (lower, upper)
def part_meet(a, b):
    return (tuple(map(max, a[0], b[0])), tuple(map(min, a[1], b[1])))


def part_move(p, v):
    return (tuple(lev_move(x, d) for x, d in zip(p[0], v)), tuple(lev_move(x, d) for x, d in zip(p[1], v)))


def part_value(p) -> int:
    vals = [x[1] for x in p[0] + p[1] if x[0] == 1]
    if len(vals) != 1:
        raise ValueError(f"malformed wall: {p}")
    return vals[0]


def walls_meet(a, b):
    return [[part_meet(p, q) for p, q in zip(pa, qa)] for pa, qa in zip(a, b)]


def walls_move(w, v):
    return [[part_move(p, v) for p in ax] for ax in w]


# ---- plan keys with the reference's orders ---------------------------------
@dataclass(frozen=True)
class IRank:
    v: Vec
    reverse: bool = True

    def key(self):
        return tuple(reversed(self.v)) if self.reverse else self.v

    def __lt__(self, o):
        return self.key() < o.key()


@dataclass(frozen=True)
class Rsc:
    """ResourceT: static (state array `name`) or OM node `nid`, with an
    optional payload (an IRank, or a (source, destination) IRank pair)."""
    static: bool
    name: object            # state name (static) or node id
    extra: object = None

    def key(self):
        if self.extra is None:
            ek = ()
        elif isinstance(self.extra, IRank):
            ek = self.extra.key()
        else:
            ek = tuple(x.key() for x in self.extra)
        return (0 if self.static else 1, self.name, ek)

    def __lt__(self, o):
        return self.key() < o.key()


@dataclass(frozen=True)
class RidgeID:
    dmpi: Vec
    delta: Rsc              # static, or OM node with (irSrc, irDest)

    def key(self):
        return (self.dmpi, self.delta.key())

    def __lt__(self, o):
        return self.key() < o.key()

    def needs_mpi(self) -> bool:
        return any(self.dmpi)


@dataclass(frozen=True)
class FacetID:
    dmpi: Vec
    src: IRank
    dest: IRank

    def key(self):
        return (self.dmpi, self.src.key(), self.dest.key())

    def __lt__(self, o):
        return self.key() < o.key()


@dataclass
class Plan:
    dim: int
    inverted: bool
    iranks: List[IRank]
    node_ids: List[int]
    region: Dict[Tuple[IRank, int], Box]          # (IRank, node) -> box
    array_alloc: Dict[Rsc, Box]                   # Rsc(static) / Rsc(node, IRank) -> box
    ridge_alloc: Dict[RidgeID, Box]
    facet_of: Dict[RidgeID, FacetID]
    facets: Dict[FacetID, List[RidgeID]]
    program: List[tuple]                          # ('wait', f) ('unstage', r) ('compute', (ir, nid), rsc) ('free', rsc) ('stage', r) ('sendrecv', f)
    system_offset: Vec
    sharing: Dict[Rsc, int]
    shared_extent: Box
    facet_tag: Dict[FacetID, int]
    support: Dict[Tuple[IRank, int], Dict[Rsc, Box]] = field(default_factory=dict)


def make_plan(step: Dict[int, object], dim: int, axes: List[str], intra: List[int],
              initial_walls: Dict[str, List[int]], state_names: List[str], inverted: bool = False,
              options: List[str] = ()) -> Plan:
    """makePlan (Cut.hs:142-160, 226-596) for one rank's step graph `step`
    (MM nodes; micro instructions as tuples)."""
    reverse_cmp = "irank-order-f" not in options
    zero = (0,) * dim

    # initialWalls (Cut.hs:182-214): per axis [0, ws..., intra]; the first and
    # last walls bound from below, the inner ones from above (swapped when
    # inverted); inverted walls are mirrored
    def wall(ax: int, ascending: bool, n: int):
        lo, hi = [TOP] * dim, [BOT] * dim
        if ascending != inverted:
            lo[ax] = lev(n)
        else:
            hi[ax] = lev(n)
        return (tuple(lo), tuple(hi))

    walls0 = []
    for a, name in enumerate(axes):
        ws = initial_walls.get(name)
        if not ws:
            raise ValueError(f"initial_walls needs at least one wall for axis {name}")
        ws2 = list(reversed([intra[a] - w for w in ws])) if inverted else list(ws)
        walls0.append([wall(a, True, 0)] + [wall(a, False, w) for w in ws2] + [wall(a, True, intra[a])])

    node_ids = sorted(step)
    wall_map = {}
    for nid in node_ids:
        acc = None
        for mn in step[nid].micro.values():
            inst = mn.inst
            if inst[0] == "LoadCursorStatic":
                w = walls_move(walls0, tuple(-c for c in inst[1]))
            elif inst[0] == "LoadCursor":
                w = walls_move(wall_map[inst[2]], tuple(-c for c in inst[1]))
            else:
                continue
            acc = w if acc is None else walls_meet(acc, w)
        if acc is None:
            raise ValueError(f"manifest node {nid} reads no grid: no walls")
        wall_map[nid] = acc

    def is_store(nid):
        n = step[nid]
        return n.micro[n.root()].inst[0] == "Store"

    def walls_key(w):
        return tuple(tuple((p[0], p[1]) for p in ax) for ax in w)

    static_walls = [wall_map[n] for n in node_ids if is_store(n)]
    consensus = min(static_walls, key=walls_key)
    wall_map2 = {n: (consensus if is_store(n) else w) for n, w in wall_map.items()}
    system_offset = tuple(part_value(consensus[a][0]) for a in range(dim))
    wall_vals = {n: [[part_value(p) for p in ax] for ax in w] for n, w in wall_map2.items()}

    # IRanks (Cut.hs:275-283)
    iranks = sorted(IRank(tuple(v), reverse_cmp)
                    for v in itertools.product(*[range(len(ax) - 1) for ax in walls0]))
    if inverted:
        iranks = list(reversed(iranks))

    def box_at(ir: IRank, nid: int, mpir: Vec = zero) -> Box:
        wv = wall_vals[nid]
        b = Box(tuple(wv[a][ir.v[a]] for a in range(dim)), tuple(wv[a][ir.v[a] + 1] for a in range(dim)))
        return b.move(tuple(m * n for m, n in zip(mpir, intra)))

    mpi_ranks = list(itertools.product(*[(-1, 0, 1)] * dim))
    mpi_box0 = Box(zero, tuple(intra))

    # supportMap (Cut.hs:335-351)
    support: Dict[Tuple[IRank, int], Dict[Rsc, Box]] = {}
    for ir in iranks:
        for nid in node_ids:
            b0 = box_at(ir, nid)
            m: Dict[Rsc, Box] = {}
            for mn in step[nid].micro.values():
                inst = mn.inst
                if inst[0] == "LoadCursorStatic":
                    r, b = Rsc(True, inst[2]), b0.move(inst[1])
                elif inst[0] == "LoadCursor":
                    r, b = Rsc(False, inst[2]), b0.move(inst[1])
                else:
                    continue
                m[r] = m[r].hull(b) if r in m else b
            support[(ir, nid)] = m

    sources = {nid: [(mpir, ir, box_at(ir, nid, mpir)) for mpir in mpi_ranks for ir in iranks] for nid in node_ids}
    static_sources = [(mpir, mpi_box0.move(tuple(m * n for m, n in zip(mpir, intra)))) for mpir in mpi_ranks]

    def locate(ir_dest: IRank, rsc: Rsc, b0: Box):
        out = []
        if not rsc.static:
            for mpir, ir, b1 in sources[rsc.name]:
                if mpir == zero and ir == ir_dest:
                    continue
                b01 = b0.meet(b1)
                if b01.volume() > 0:
                    out.append((RidgeID(mpir, Rsc(False, rsc.name, (ir, ir_dest))), b01))
        else:
            for mpir, b1 in static_sources:
                if mpir == zero:
                    continue
                b01 = b0.meet(b1)
                if b01.volume() > 0:
                    out.append((RidgeID(mpir, Rsc(True, rsc.name)), b01))
        return out

    ridge_request: Dict[Tuple[IRank, int], List[Tuple[RidgeID, Box]]] = {}
    for key, m in support.items():
        ridge_request[key] = [rb for r in sorted(m) for rb in locate(key[0], r, m[r])]
    first_needed: Dict[RidgeID, IRank] = {}
    for ir in iranks:
        for nid in node_ids:
            for rid, _ in ridge_request[(ir, nid)]:
                first_needed.setdefault(rid, ir)
    ridges: Dict[RidgeID, Box] = {}
    for ir in iranks:
        for nid in node_ids:
            for rid, b in ridge_request[(ir, nid)]:
                ridges[rid] = ridges[rid].hull(b) if rid in ridges else b
    ridge_keys = sorted(ridges)

    # ridgeProvide (Cut.hs:423-430): who fills each ridge
    provide: Dict[Rsc, List[RidgeID]] = {}
    for r in ridge_keys:
        k = Rsc(True, r.delta.name) if r.delta.static else Rsc(False, r.delta.name, r.delta.extra[0])
        provide.setdefault(k, []).append(r)

    facet_of: Dict[RidgeID, FacetID] = {}
    for r in ridge_keys:
        if r.needs_mpi():
            src = iranks[0] if r.delta.static else r.delta.extra[0]
            facet_of[r] = FacetID(r.dmpi, src, first_needed[r])
    facets: Dict[FacetID, List[RidgeID]] = {}
    for r in ridge_keys:
        if r in facet_of:
            facets.setdefault(facet_of[r], []).append(r)
    facet_keys = sorted(facets)

    # distributed program (Cut.hs:444-490)
    prog: List[tuple] = []
    issued = set()

    def insert(inst):
        issued.add(inst)
        prog.append(inst)

    def insert_once(inst):
        if inst not in issued:
            insert(inst)

    for sn in sorted(state_names):
        for r in provide.get(Rsc(True, sn), []):
            insert_once(("stage", r))
    for ir in iranks:
        for f in facet_keys:
            if f.dest == ir:
                insert(("wait", f))
        for nid in node_ids:
            for r, _ in ridge_request[(ir, nid)]:
                insert_once(("unstage", r))
            n = step[nid]
            root = n.micro[n.root()].inst
            tail = Rsc(True, root[1]) if root[0] == "Store" else Rsc(False, nid, ir)
            insert(("compute", (ir, nid), tail))
            for r in provide.get(Rsc(False, nid, ir), []):
                insert_once(("stage", r))
        for f in facet_keys:
            if f.src == ir:
                insert(("sendrecv", f))

    # allocations (Cut.hs:493-507)
    alloc: Dict[Rsc, Box] = {}

    def grow(k, b):
        alloc[k] = alloc[k].hull(b) if k in alloc else b

    for sn in state_names:
        grow(Rsc(True, sn), mpi_box0)
    for ir in iranks:
        for nid in node_ids:
            grow(Rsc(False, nid, ir), box_at(ir, nid))
    for r in ridge_keys:
        if r.delta.static:
            grow(Rsc(True, r.delta.name), ridges[r])
        else:
            grow(Rsc(False, r.delta.name, r.delta.extra[1]), ridges[r])

    # FreeResource after the last computation that reads a region's array
    # (Cut.hs:512-528), then buffer sharing by simulation (:530-545)
    last_used: Dict[Rsc, float] = {}
    for ln, inst in enumerate(prog):
        if inst[0] == "compute":
            ir, nid = inst[1]
            for r in support[(ir, nid)]:
                if not r.static:
                    k = Rsc(False, r.name, ir)
                    last_used[k] = max(last_used.get(k, -1.0), float(ln))
    numbered = [(float(i), inst) for i, inst in enumerate(prog)] + \
        [(ln + 0.5, ("free", k)) for k, ln in last_used.items()]
    numbered.sort(key=lambda x: x[0])
    prog1 = [inst for _, inst in numbered]
    sharing: Dict[Rsc, int] = {}
    free_ids: List[int] = []
    next_id = 0
    for inst in prog1:
        if inst[0] == "free":
            if inst[1] in sharing:
                free_ids.insert(0, sharing[inst[1]])
        elif inst[0] == "compute" and not inst[2].static:
            if inst[2] not in sharing:
                if free_ids:
                    sharing[inst[2]] = free_ids.pop(0)
                else:
                    sharing[inst[2]] = next_id
                    next_id += 1
    om_boxes = [Box(zero, b.shape()) for k, b in alloc.items() if not k.static]
    largest = max(om_boxes, key=lambda b: b.key()) if om_boxes else Box(zero, zero)

    # optimizeCommunicationsOrder (Cut.hs:599-616): runs of send-receives in
    # the order of their waits
    wait_pos = {inst[1]: i for i, inst in enumerate(prog1) if inst[0] == "wait"}
    prog2: List[tuple] = []
    run: List[tuple] = []
    for inst in prog1 + [None]:
        if inst is not None and inst[0] == "sendrecv":
            run.append(inst)
            continue
        if run:
            prog2.extend(sorted(run, key=lambda x: (x[1] in wait_pos, wait_pos.get(x[1], 0))))
            run = []
        if inst is not None:
            prog2.append(inst)

    # a facet must be sent (by the neighbour rank, which runs the same
    # program) before it is waited for: the reference emits MPI_Wait on the
    # facet's request at the start of the destination region and the
    # Irecv/Isend pair after the source region, so a facet whose source region
    # comes later (or is the destination itself) would be read stale.  Such a
    # plan arises when the walls leave regions narrower than the stencil
    # reach of the (T-fold) step; reject it.
    pos = {}
    for i, inst in enumerate(prog2):
        if inst[0] in ("wait", "sendrecv"):
            pos.setdefault(inst[1], {})[inst[0]] = i
    for f, where in pos.items():
        if where.get("sendrecv", len(prog2)) > where.get("wait", len(prog2)):
            raise ValueError(f"walls {initial_walls} give regions narrower than the stencil reach: facet "
                             f"{f.dmpi} from region {f.src.v} is waited for before it is sent")

    region = {(ir, nid): box_at(ir, nid) for ir in iranks for nid in node_ids}
    return Plan(dim, inverted, iranks, node_ids, region, alloc, ridges, facet_of, facets, prog2,
                system_offset, sharing, largest, {f: i for i, f in enumerate(facet_keys)}, support)
