"""The coverage argument behind the dry test of dry-tile skipping, checked on the CPU.

dry_region_of (moses_2pdf_amd/csrc/m2p_tiles.h) lists, for the region a stage tile or a
CFL box reads, the row-segment flags of its interior cells and the corner ghost cells it
holds; no other ghost cell is tested.  That is sound because the ghost fill makes every
non-corner ghost cell a 0.5-weighted sum of interior cells two deep along the boundary in
the same row or column (momenta 0).  This test restates the item enumeration and checks,
for every tile of every compiled tile height and every CFL box of many grid shapes, that
(a) every item is in bounds, and (b) every cell of the region is either listed (corner),
covered by a listed flag (interior), or a ghost cell whose source cells are covered.  It also
checks the bounds dry_flags_wet3 (the same flag items, walked row by row) relies on."""
MD = 3


def region_items(nx, ny, ra, rb, ca, cb):
    nxd, nyd = nx + 2 * MD, ny + 2 * MD
    ra, rb, ca, cb = max(ra, 0), min(rb, nyd - 1), max(ca, 0), min(cb, nxd - 1)
    xmax, ymax = nx + MD - 1, ny + MD - 1
    ia, ib = max(ra, MD), min(rb, ymax)
    sa = (max(ca, MD) - MD) // 63
    ns = (min(cb, xmax) - MD) // 63 - sa + 1
    rt, nrt = ra, max(0, min(rb, MD - 1) - ra + 1)
    rbo = max(ra, ymax + 1)
    nrg = nrt + max(0, rb - rbo + 1)
    cl, ncl = ca, max(0, min(cb, MD - 1) - ca + 1)
    cr = max(ca, xmax + 1)
    ncg = ncl + max(0, cb - cr + 1)
    nseg = (nx + 62) // 63
    flags, corners = set(), set()
    for e in range((ib - ia + 1) * ns):
        r, s = ia + e // ns, sa + e % ns
        assert MD <= r <= ymax and 0 <= s < nseg
        flags.add((r, s))
    # what dry_flags_wet3 walks row by row: rows [ia, ib], segments [sa, sa + ns)
    assert flags == {(r, sa + j) for r in range(ia, ib + 1) for j in range(ns)}
    for e in range(nrg * ncg):
        ir, ic = e // ncg, e % ncg
        r = rt + ir if ir < nrt else rbo + (ir - nrt)
        c = cl + ic if ic < ncl else cr + (ic - ncl)
        assert 0 <= r < nyd and 0 <= c < nxd
        assert (r < MD or r > ymax) and (c < MD or c > xmax)
        corners.add((r, c))
    return (ra, rb, ca, cb), flags, corners


def check_region(nx, ny, ra, rb, ca, cb, walked=False):
    (ra, rb, ca, cb), flags, corners = region_items(nx, ny, ra, rb, ca, cb)
    if walked:   # dry_flags_wet3's bounds (stage tiles: TY + 4 rows, CFL boxes 7): 3 segments
        assert len({s for _, s in flags}) <= 3 and rb - ra + 1 <= 19
    xmax, ymax = nx + MD - 1, ny + MD - 1
    cov = lambda r, c: (r, (c - MD) // 63) in flags
    for r in range(ra, rb + 1):
        for c in range(ca, cb + 1):
            gr, gc = r < MD or r > ymax, c < MD or c > xmax
            if not gr and not gc:
                assert cov(r, c)
            elif gr and gc:
                assert (r, c) in corners
            elif gc:
                src = [(r, MD), (r, MD + 1)] if c < MD else [(r, xmax), (r, xmax - 1)]
                assert all(cov(*s) for s in src)
            else:
                src = [(MD, c), (MD + 1, c)] if r < MD else [(ymax, c), (ymax - 1, c)]
                assert all(cov(*s) for s in src)


def test_dry_regions_cover_their_stencils():
    shapes = [(2, 2), (3, 5), (62, 7), (63, 63), (64, 9), (65, 130), (126, 40), (127, 200),
              (200, 3), (257, 129)]
    n = 0
    for nx, ny in shapes:
        for TY in (3, 7, 15):
            for tx in range((nx + 62) // 63):
                for ty in range((ny + TY - 1) // TY):
                    c0, r0 = MD + 63 * tx, MD + TY * ty
                    check_region(nx, ny, r0 - 2, r0 + TY + 1, c0 - 2, c0 + 64, True)
                    n += 1
        for bx in range((nx + 63) // 64):
            for by in range((ny + 3) // 4):
                clo, rlo = MD + 64 * bx, MD + 4 * by
                check_region(nx, ny, rlo - 1, rlo + 5, clo - 1, clo + 65, True)
                n += 1
    assert n > 1000


def ring_face_pos(nx, ny, t):
    nxd = nx + 2 * MD
    f, d = t >> 1, t & 1
    w = nxd - 3
    nrow, ncol = 4 * w, 4 * ny
    if f >= nrow + ncol:
        return None
    if f < nrow:
        k = f // w
        r = 1 + k if k < 2 else ny + MD + (k - 2)
        c = 1 + f % w
    else:
        q = f - nrow
        k = q // ny
        r = MD + q % ny
        c = 1 + k if k < 2 else nx + MD + (k - 2)
    return r, c, d


def ring_face_reads(nx, ny, r, c, d):
    """Cells final_states_edge reads for face (r, c, dir d): Interfaces2's row and column
    overrides, then base_states' 4-cell stencil."""
    rows = [MD, MD + 1] if r <= 2 else [ny + MD - 1, ny + MD - 2] if r >= ny + MD else [r]
    cols = [MD, MD + 1] if c <= 2 else [nx + MD - 1, nx + MD - 2] if c >= nx + MD else [c]
    out = set()
    for rr in rows:
        for cc in cols:
            for k in range(-1, 3):
                out.add((rr, cc + k) if d == 0 else (rr + k, cc))
    return out


def test_ring_dry_regions_cover_their_faces():
    for nx, ny in [(2, 2), (5, 3), (40, 70), (63, 63), (130, 64), (257, 129)]:
        nthreads = 2 * (4 * (nx + 2 * MD - 3) + 4 * ny)
        for b in range((nthreads + 63) // 64):
            faces = [ring_face_pos(nx, ny, t) for t in range(64 * b, 64 * b + 64)]
            faces = [f for f in faces if f]
            ra = rb = ca = cb = None
            for r, c, d in faces:
                gr, gc = r <= 2 or r >= ny + MD, c <= 2 or c >= nx + MD
                lo_r = (MD if r <= 2 else ny + MD - 2) if gr else r
                hi_r = (MD + 1 if r <= 2 else ny + MD - 1) if gr else r
                lo_c = (MD if c <= 2 else nx + MD - 2) if gc else c
                hi_c = (MD + 1 if c <= 2 else nx + MD - 1) if gc else c
                ra = lo_r if ra is None else min(ra, lo_r)
                rb = hi_r if rb is None else max(rb, hi_r)
                ca = lo_c if ca is None else min(ca, lo_c)
                cb = hi_c if cb is None else max(cb, hi_c)
            reg = (ra - 1, rb + 2, ca - 1, cb + 2)
            for r, c, d in faces:
                for (rr, cc) in ring_face_reads(nx, ny, r, c, d):
                    assert reg[0] <= rr <= reg[1] and reg[2] <= cc <= reg[3]
            check_region(nx, ny, *reg)
