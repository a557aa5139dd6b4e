"""Known answers that pin the oracle in 2D and 3D where the golden fixtures
(2D advection_2d only) leave it open.  Each check restates a result of the
reference's own tests or of the mathematics, not the oracle's formulas:

- projection/prediction round trip on a two-level mesh: cell averages of a
  separable polynomial of per-axis degree 2s are reproduced by projection then
  prediction <s,false> and <s,true> (tests/test_projection_prediction_roundtrip.cpp:
  139-197, 214-244, tolerances :375-446);
- update_ghost_mr on an adapted 3D mesh reproduces such averages at every
  interior projection and prediction ghost;
- the multiresolution details vanish on per-axis quadratics
  (mr/operators.hpp:360-533 predicts them exactly) and not on cubics;
- Dirichlet<1> ghost values on uniform and adapted meshes
  (tests/test_bc_ghost_values.cpp:120-175, 540-578, 707-726);
- corner ghosts: the reflection of the field about the domain corner
  (tests/test_bc_ghost_values.cpp:278-358);
- the upwind step on a uniform 3D mesh equals the textbook donor-cell update
  with Dirichlet-0 ghosts, for velocities of both signs.
"""
import numpy as np
import pytest

from initial_conditions import centers
from oracle_lib import CELLS, REFERENCE, OracleSim


def antideriv_quadratic(t):
    # A(t) = t + t^2 + t^3, the antiderivative of 1 + 2t + 3t^2
    # (tests/test_projection_prediction_roundtrip.cpp:226-240)
    t2 = t * t
    return t + t2 + t2 * t


def antideriv_cubic(t):
    # antiderivative of 1 + 2t + 3t^2 + 4t^3: per-axis degree 3, beyond what
    # the s = 1 prediction reproduces
    t2 = t * t
    return t + t2 + t2 * t + t2 * t2


def cell_average(rows, dim, antiderivs):
    """Cell averages of prod_d p_d(x_d) (cell_avg_nd of the round-trip test):
    prod_d (A_d(c + h/2) - A_d(c - h/2)) / h."""
    h = np.ldexp(1.0, -rows[:, 0].astype(np.int64))
    c = centers(rows)
    out = np.ones(len(rows))
    for d in range(dim):
        A = antiderivs[d]
        out *= (A(c[:, d] + h / 2.0) - A(c[:, d] - h / 2.0)) / h
    return out


@pytest.mark.parametrize("dim,s,L,tol", [(2, 0, 4, 1e-14), (3, 0, 4, 1e-14), (2, 1, 4, 1e-13), (3, 1, 4, 1e-13)])
def test_projection_prediction_roundtrip(dim, s, L, tol):
    """roundtrip_error (tests/test_projection_prediction_roundtrip.cpp:139-197):
    every reference slot of a two-level mesh (L-1, L) holds the exact average;
    projection L -> L-1, then the fine cells are poisoned and predicted from
    L-1, by <s,false> and by <s,true>."""
    sim = OracleSim(dim, L - 1, L, pred_s=s)
    ref = sim.cells(REFERENCE)
    if s == 0:
        exact = np.full(len(ref), 3.14)
    else:
        exact = cell_average(ref, dim, [antideriv_quadratic] * dim)
    st = np.full(sim.nb_cells(REFERENCE), np.nan)
    st[ref[:, 4]] = exact
    sim.set_storage(st)
    sim.apply_level_op(0, L)
    fine = sim.cells(CELLS)
    assert np.all(fine[:, 0] == L)
    idx_of = {int(i): k for k, i in enumerate(ref[:, 4])}
    fine_exact = np.array([exact[idx_of[int(i)]] for i in fine[:, 4]])
    errors = []
    for op in (1, 2):
        st = sim.storage()
        st[fine[:, 4]] = np.nan
        sim.set_storage(st)
        sim.apply_level_op(op, L)
        got = sim.storage()[fine[:, 4]]
        assert not np.isnan(got).any(), f"op {op}: a fine cell was not predicted"
        errors.append(np.max(np.abs(got - fine_exact)))
    assert max(errors) <= tol, errors


def adapted_mesh_cells(dim, lmin=2, lmax=5):
    """An adapted mesh with leaves on several levels inside the box."""
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    c = centers(o.cells())
    o.set_values(np.exp(-np.sum((c[:, :dim] - 0.5) ** 2, axis=1) / (2 * 0.08 ** 2)))
    o.set_dirichlet(0.0)
    o.adapt(1e-2, 2.0)
    cells = o.cells()
    assert len(np.unique(cells[:, 0])) >= 3
    return cells


@pytest.mark.parametrize("dim", [2, 3])
def test_ghost_update_reproduces_quadratic_averages(dim):
    """update_ghost_mr on an adapted mesh from exact leaf averages of a
    per-axis quadratic: every projection ghost (average of exact children) and
    every prediction ghost whose stencils stay inside the box holds the exact
    average too.  Outer ghosts start as NaN and no boundary condition is set,
    so a value computed from outside the box shows as NaN."""
    lmin, lmax = 2, 5
    cells = adapted_mesh_cells(dim, lmin, lmax)
    sim = OracleSim(dim, lmin, lmax, pred_s=1, cells=cells)
    ref = sim.cells(REFERENCE)
    exact = cell_average(ref, dim, [antideriv_quadratic] * dim)
    leaves = sim.cells(CELLS)
    pos = {int(i): k for k, i in enumerate(ref[:, 4])}
    st = np.full(sim.nb_cells(REFERENCE), np.nan)
    st[leaves[:, 4]] = exact[[pos[int(i)] for i in leaves[:, 4]]]
    sim.set_storage(st)
    sim.update_ghost_mr()
    out = sim.storage()[ref[:, 4]]
    c = centers(ref)
    is_leaf = np.isin(ref[:, 4], leaves[:, 4])
    region = np.all((c[:, :dim] > 0.3) & (c[:, :dim] < 0.7), axis=1) & (ref[:, 0] >= 3) & ~is_leaf
    assert region.sum() > (150 if dim == 3 else 40), region.sum()
    assert not np.isnan(out[region]).any()
    assert np.max(np.abs(out[region] - exact[region])) <= 1e-13
    # both kinds of ghost are in the region: coarser than a leaf (projection)
    # and finer than a leaf (prediction)
    by_level = {}
    for r in leaves:
        by_level.setdefault(int(r[0]), set()).add(tuple(r[1:4]))
    coarser = finer = 0
    for r in ref[region]:
        l, x, y, z = (int(v) for v in r[:4])
        parent = (x >> 1, y >> 1, z >> 1 if dim == 3 else 0)
        if parent in by_level.get(l - 1, set()) or any(
                (x >> k, y >> k, (z >> k) if dim == 3 else 0) in by_level.get(l - k, set()) for k in range(2, l - lmin + 1)):
            finer += 1
        else:
            coarser += 1
    assert coarser > 0 and finer > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_details_vanish_on_quadratics(dim):
    """compute_detail_op (mr/operators.hpp:226-533): the detail is the fine
    value minus its prediction from the coarse stencil, so it vanishes for
    per-axis quadratics and not for a cubic in any one axis."""
    lmin, lmax = 2, 5
    cells = adapted_mesh_cells(dim, lmin, lmax)

    def details(antiderivs):
        sim = OracleSim(dim, lmin, lmax, pred_s=1, cells=cells)
        ref = sim.cells(REFERENCE)
        leaves = sim.cells(CELLS)
        pos = {int(i): k for k, i in enumerate(ref[:, 4])}
        exact = cell_average(ref, dim, antiderivs)
        st = np.full(sim.nb_cells(REFERENCE), np.nan)
        st[leaves[:, 4]] = exact[[pos[int(i)] for i in leaves[:, 4]]]
        sim.set_storage(st)
        sim.harten(0, 1e-3, 2.0)
        _, d = sim.last_tags()
        c = centers(ref)
        region = np.all((c[:, :dim] > 0.3) & (c[:, :dim] < 0.7), axis=1) & (ref[:, 0] >= 3)
        return d[ref[region, 4]]

    d2 = details([antideriv_quadratic] * dim)
    assert not np.isnan(d2).any()
    assert np.max(np.abs(d2)) <= 1e-13
    for axis in range(dim):
        anti = [antideriv_quadratic] * dim
        anti[axis] = antideriv_cubic
        d3 = details(anti)
        assert np.sum(np.abs(d3) > 1e-9) > 20, f"cubic along axis {axis} left no detail"


def cartesian_directions(dim):
    # for_each_cartesian_direction (stencil.hpp:300-316): +e0, -e0, +e1, ...
    for d in range(dim):
        for sgn in (1, -1):
            e = [0] * dim
            e[d] = sgn
            yield d, sgn, e


def check_dirichlet1(cells, dim, lmin, lmax):
    """run_dirichlet<dim, 1>(mesh, false) (tests/test_bc_ghost_values.cpp:130-175,
    check_ghosts :104-124): per direction a field f = 1 + 3 x_n, the boundary
    value f on the face, apply_field_bc(u, direction); every first-layer ghost
    beside a boundary leaf holds f(ghost centre)."""
    checked = 0
    for axis, sgn, e in cartesian_directions(dim):
        sim = OracleSim(dim, lmin, lmax, pred_s=1, cells=cells)
        leaves = sim.cells(CELLS)
        f = lambda rows: 1.0 + 3.0 * centers(rows)[:, axis]
        sim.set_values(f(leaves))
        sim.set_dirichlet(1.0 + 3.0 * (1.0 if sgn > 0 else 0.0))
        sim.apply_field_bc_direction(e)
        ref = sim.cells(REFERENCE)
        pos = {tuple(int(v) for v in r[:4]): int(r[4]) for r in ref}
        st = sim.storage()
        for r in leaves:
            l = int(r[0])
            g = [int(r[1]), int(r[2]), int(r[3])]
            g[axis] += sgn
            if 0 <= g[axis] < (1 << l):
                continue
            key = (l, g[0], g[1], g[2])
            assert key in pos, f"boundary ghost {key} missing from the reference mesh"
            ghost = np.array([[l, g[0], g[1], g[2], 0]])
            assert abs(st[pos[key]] - f(ghost)[0]) <= 1e-11, (axis, sgn, key)
            checked += 1
    return checked


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet1_uniform(dim):
    # dirichlet1_constant_{2d,3d} (tests/test_bc_ghost_values.cpp:549-559): level 4
    sim = OracleSim(dim, 4, 4, pred_s=1)
    assert check_dirichlet1(sim.cells(), dim, 4, 4) == 2 * dim * 16 ** (dim - 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet1_adapted(dim):
    """dirichlet1_adapted_{2d,3d} (tests/test_bc_ghost_values.cpp:686-726): an
    indicator ball of radius 0.3 at the origin corner, adapted with eps 1e-4
    on L2..5, puts boundary leaves of several levels on the '-' faces."""
    o = OracleSim(dim, 2, 5, pred_s=1)
    c = centers(o.cells())
    o.set_values(np.where(np.sqrt(np.sum(c[:, :dim] ** 2, axis=1)) < 0.3, 1.0, 0.0))
    o.set_dirichlet(0.0)
    o.adapt(1e-4, 1.0)
    cells = o.cells()
    # adapted_boundary_crosses_levels: boundary leaves toward -x on > 1 level
    assert len(np.unique(cells[cells[:, 1] == 0, 0])) > 1
    assert check_dirichlet1(cells, dim, 2, 5) > 0


def corner_oracle(c, dim, f):
    """corner_oracle (tests/test_bc_ghost_values.cpp:278-306): the reflection
    of the ghost centre about the domain corner, along the first outside axis."""
    out = np.zeros(len(c))
    for k, p in enumerate(c):
        outward = [(-1 if p[d] < 0 else (1 if p[d] > 1 else 0)) for d in range(dim)]
        first = next(d for d in range(dim) if outward[d] != 0)
        mag = abs(p[first] - (0.0 if outward[first] < 0 else 1.0))
        q = p[:dim].copy()
        for d in range(dim):
            if outward[d] != 0:
                q[d] = (0.0 if outward[d] < 0 else 1.0) - mag * outward[d]
        out[k] = f(q[None, :])[0]
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_corner_ghosts_reflect_the_field(dim):
    """run_corners (tests/test_bc_ghost_values.cpp:308-358), ghost width 1 on a
    uniform level-3 mesh: every ghost outside the box along two or more axes
    (the 2D corners, the 3D edges and vertices) equals the field at the
    reflection of its centre."""
    f = lambda q: 1.0 + 0.5 * q[:, 0] * q[:, 0] + sum((d + 1) * q[:, d] for d in range(dim))
    sim = OracleSim(dim, 3, 3, pred_s=1)
    leaves = sim.cells()
    sim.set_values(f(centers(leaves)))
    sim.update_ghost_mr()
    ref = sim.cells(REFERENCE)
    ref3 = ref[ref[:, 0] == 3]
    c = centers(ref3)
    nout = np.sum((c[:, :dim] < 0) | (c[:, :dim] > 1), axis=1)
    corner = ref3[nout >= 2]
    expected_count = 4 if dim == 2 else 8 + 12 * 8
    assert len(corner) == expected_count
    got = sim.storage()[corner[:, 4]]
    assert np.max(np.abs(got - corner_oracle(centers(corner), dim, f))) <= 1e-11


@pytest.mark.parametrize("a", [(1.0, 1.0, 1.0), (0.7, -0.3, 0.5), (-0.25, 0.6, -0.9)])
def test_upwind_3d_uniform_closed_form(a):
    """unp1 = u - dt * upwind(a, u) on a uniform 3D mesh is the donor-cell
    update u - dt/dx sum_d (F_{d,+} - F_{d,-}) with F = a u_upwind-side, and
    Dirichlet 0 ghosts -u (bc/dirichlet.hpp:29)."""
    n, L = 16, 4
    sim = OracleSim(3, L, L, pred_s=1)
    cells = sim.cells()
    rng = np.random.default_rng(7)
    u = rng.random(len(cells))
    sim.set_values(u)
    sim.set_dirichlet(0.0)
    dt = 0.4 / n / 3
    sim.upwind_step(np.array(a), dt)
    g = np.zeros((n + 2, n + 2, n + 2))   # [z, y, x] with one ghost layer
    g[cells[:, 3] + 1, cells[:, 2] + 1, cells[:, 1] + 1] = u
    for ax in range(3):
        lo = [slice(1, n + 1)] * 3
        hi = [slice(1, n + 1)] * 3
        src_lo, src_hi = list(lo), list(hi)
        lo[ax], src_lo[ax] = 0, 1
        hi[ax], src_hi[ax] = n + 1, n
        g[tuple(lo)] = -g[tuple(src_lo)]
        g[tuple(hi)] = -g[tuple(src_hi)]
    inner = g[1:-1, 1:-1, 1:-1]
    total = np.zeros_like(inner)
    for d, ax in ((0, 2), (1, 1), (2, 0)):   # velocity component d acts along array axis ax
        minus = np.take(g, np.arange(0, n), axis=ax)[tuple(slice(1, -1) if k != ax else slice(None) for k in range(3))]
        plus = np.take(g, np.arange(2, n + 2), axis=ax)[tuple(slice(1, -1) if k != ax else slice(None) for k in range(3))]
        if a[d] >= 0:
            total += a[d] * (inner - minus)
        else:
            total += a[d] * (plus - inner)
    expect = inner - dt * n * total
    got = np.zeros_like(inner)
    got[cells[:, 3], cells[:, 2], cells[:, 1]] = sim.values()
    assert np.max(np.abs(got - expect)) <= 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_vertex_ghosts_cascade_from_the_corner_leaf(dim):
    """project_corner_below (algorithm/update_outer_ghost.hpp:264-336): the
    vertex ghost at every level below the leaf that touches a vertex of the box
    holds that leaf's value (the leaf's own vertex ghost is a copy of it, and
    each coarser vertex ghost copies its corner-most child).  On the adapted
    mesh of a ball at the origin the (0, ..., 0) vertex carries a max-level
    leaf while the other vertices are coarse."""
    f = lambda q: 1.0 + 0.5 * q[:, 0] * q[:, 0] + sum((d + 1) * q[:, d] for d in range(dim))
    o = OracleSim(dim, 2, 5, pred_s=1)
    c = centers(o.cells())
    o.set_values(np.where(np.sqrt(np.sum(c[:, :dim] ** 2, axis=1)) < 0.3, 1.0, 0.0))
    o.set_dirichlet(0.0)
    o.adapt(1e-4, 1.0)
    sim = OracleSim(dim, 2, 5, pred_s=1, cells=o.cells())
    leaves = sim.cells(CELLS)
    sim.set_values(f(centers(leaves)))
    sim.set_dirichlet(0.0)
    sim.update_ghost_mr()
    st = sim.storage()
    ref = sim.cells(REFERENCE)
    pos = {tuple(int(v) for v in r[:4]): int(r[4]) for r in ref}
    checked = 0
    deepest = 0
    for corner in range(1 << dim):
        side = [(corner >> d) & 1 for d in range(dim)]   # 0: the '-' face, 1: the '+' face
        # the leaf touching this vertex
        leaf = None
        for r in leaves:
            l = int(r[0])
            if all(int(r[1 + d]) == (0 if side[d] == 0 else (1 << l) - 1) for d in range(dim)):
                leaf = r
        assert leaf is not None
        lf = int(leaf[0])
        value = st[int(leaf[4])]
        for l in range(0, lf + 1):
            g = [(-1 if side[d] == 0 else (1 << l)) for d in range(dim)] + [0] * (3 - dim)
            key = (l, g[0], g[1], g[2])
            if key in pos:
                assert st[pos[key]] == value, (corner, l)
                checked += 1
                deepest = max(deepest, lf - l)
    assert checked >= 2 ** dim and deepest >= 2


def test_edge_ghosts_cascade_below_the_finest_level():
    """project_corner_below along the twelve edges of the cube
    (algorithm/update_outer_ghost.hpp:264-336): on a uniform level-4 mesh with
    min_level 2, every edge ghost of the reference mesh at levels 3 and 2
    (outside the box along exactly two axes) copies the level-4 edge ghost of
    its corner-most child -- the child next to the box along the two outside
    axes and the first along the edge -- whose value is the field at the
    reflection of that child's centre (corner_oracle)."""
    dim, L = 3, 4
    f = lambda q: 1.0 + 0.5 * q[:, 0] * q[:, 0] + q[:, 0] + 2.0 * q[:, 1] + 3.0 * q[:, 2]
    sim = OracleSim(dim, 2, L, pred_s=1)
    sim.set_values(f(centers(sim.cells())))
    sim.update_ghost_mr()
    st = sim.storage()
    ref = sim.cells(REFERENCE)
    checked = 0
    for l in (3, 2):
        d = L - l
        rows = ref[ref[:, 0] == l]
        n = 1 << l
        for r in rows:
            g = [int(v) for v in r[1:4]]
            out = [(-1 if g[k] < 0 else (1 if g[k] >= n else 0)) for k in range(3)]
            if sum(o != 0 for o in out) != 2 or any(g[k] < -1 or g[k] > n for k in range(3)):
                continue
            child = [(g[k] << d) + ((1 << d) - 1 if out[k] == -1 else 0) for k in range(3)]
            fine = np.array([[L, child[0], child[1], child[2], 0]])
            expect = corner_oracle(centers(fine), dim, f)[0]
            assert abs(st[int(r[4])] - expect) <= 1e-11, (l, g)
            checked += 1
    assert checked >= 12 * 2


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_several_fields_reduce_to_one(dim, lmin, lmax):
    """make_MRAdapt(u, v) (mr/adapt.hpp:138-146, criteria.hpp's n_comp loops):
    with v = u the first adaptation is the one of u alone (same leaves and
    values, and v carried along as u); with v = 0 every step (adaptation +
    upwind of u) is the one of u alone.  A refinement that needed every field
    instead of any, or a coarsening allowed by any field instead of all,
    fails the v = 0 case."""
    from initial_conditions import disc_2d, sphere_3d
    f = disc_2d if dim == 2 else sphere_3d
    a = np.array([1.0, 0.5, 0.25][:dim])
    dt = 0.4 / 3 * 2.0 ** -lmax
    eps = 1e-4 if dim == 2 else 1e-3
    sims = []
    for second in (None, "same", "zero"):
        o = OracleSim(dim, lmin, lmax, pred_s=1)
        o.set_poison(False)
        c = o.cells()
        o.set_values(f(c))
        o.set_dirichlet(0.0)
        if second is not None:
            o.add_field(f(c) if second == "same" else np.zeros(len(c)))
        sims.append(o)
    calls = [o.adapt(eps, 2.0) for o in sims]
    assert calls[0] == calls[1] == calls[2]
    for o in sims[1:]:
        assert np.array_equal(o.cells(), sims[0].cells())
        assert np.array_equal(o.values().view(np.uint64), sims[0].values().view(np.uint64))
    assert np.array_equal(sims[1].extra_values(0).view(np.uint64), sims[1].values().view(np.uint64))
    one, zero = sims[0], sims[2]
    for step in range(3):
        one.upwind_step(a, dt)
        zero.upwind_step(a, dt)
        assert one.adapt(eps, 2.0) == zero.adapt(eps, 2.0), step
        assert np.array_equal(one.cells(), zero.cells()), step
        assert np.array_equal(one.values().view(np.uint64), zero.values().view(np.uint64)), step
        assert not np.any(zero.extra_values(0)), step


def _adapted(dim, lmin, lmax, values, relative, eps=1e-3):
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_values(values(o.cells()))
    o.set_dirichlet(0.0)
    o.set_relative_detail(relative)
    o.adapt(eps, 2.0)
    return o.cells()


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_relative_detail_scale_invariance(dim, lmin, lmax):
    """mra_config().relative_detail(true) (mr/rel_detail.hpp:74-112) divides
    every detail by max |u| over the leaves: the adapted mesh of c * u equals
    that of u for any c > 0 (c a power of two, so the scaling is exact);
    with max |u| = 1 it is the absolute criterion's mesh, and with u scaled
    by 8 the absolute criterion refines more while the relative one does not
    move."""
    from initial_conditions import disc_2d, sphere_3d
    ic = disc_2d if dim == 2 else sphere_3d
    base = _adapted(dim, lmin, lmax, ic, True)
    assert np.array_equal(base, _adapted(dim, lmin, lmax, lambda c: 4.0 * ic(c), True))
    assert np.array_equal(base, _adapted(dim, lmin, lmax, lambda c: 0.25 * ic(c), True))
    assert np.array_equal(base, _adapted(dim, lmin, lmax, ic, False))   # max |u| = 1
    assert len(_adapted(dim, lmin, lmax, lambda c: 8.0 * ic(c), False)) > len(base)


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 6), (3, 2, 4)])
def test_refine_boundary_keeps_the_boundary_layer(dim, lmin, lmax):
    """--refine-boundary (keep_boundary_refined, mr/adapt.hpp:245-272, 341-346):
    the finest leaves within the stencil radius (1) of a face are tagged keep.
    On a zero field every detail vanishes, so without the option the mesh
    coarsens to min_level everywhere; with it the finest leaves are exactly
    the boundary layer closed over sibling groups (a parent coarsens only
    when all its children do): the cells of max_level with a coordinate in
    {0, 1} or {n-2, n-1}."""
    def run(flag):
        o = OracleSim(dim, lmin, lmax, pred_s=1)
        o.set_values(np.zeros(o.nb_cells()))
        o.set_dirichlet(0.0)
        o.set_refine_boundary(flag)
        o.adapt(1e-3, 2.0)
        return o.cells()
    plain = run(False)
    assert np.all(plain[:, 0] == lmin)
    kept = run(True)
    fine = kept[kept[:, 0] == lmax]
    n = 1 << lmax
    near = np.zeros(len(fine), dtype=bool)
    for d in range(dim):
        x = fine[:, 1 + d]
        near |= (x < 2) | (x >= n - 2)
    assert np.all(near)
    assert len(fine) == n ** dim - (n - 4) ** dim
