#!/usr/bin/env python3
"""Generates kernels_f32.hip / kernels_f64.hip: the instantiated row/column
pass kernels and the schedule table of each (precision, Nx/2, Ny).

A schedule fixes the brick shape (BY rows x BX columns), for each pass the
radix R and how many fields share an LDS exchange, and the number of columns
per column-pass block (BX or a divisor of it: narrower blocks mean more of
them, which lets a lone run overlap one block's loads with another's math).  For every size there
is one default schedule (used unless a tuned table in schedules.txt says
otherwise) and, for square grids from 256^2 to 4096^2, a few alternatives for
the autotuner (autotune.py).  Constraints checked here: at most 512 lanes per
block and at most 160 KB of LDS per block.

    python3 gen_kernel_tables.py OUTDIR   # writes the .hip files into OUTDIR

The Makefile runs it into its build directory; the generated files are not
kept in the source tree.
"""
import os
import sys

SIZES = [32, 64, 128, 256, 512, 1024, 2048, 4096]
LDS_MAX = 160 * 1024


def lds_area(L, W, R):
    return L * W + (L * W) // R + 1


def lds_row(nx2, rb, R, G, esz, pp=0):
    return ((2 if pp & 1 else 1) * G * lds_area(nx2, rb, R) + 2 * nx2) * esz


def lds_col(ny, W, R, GF, GI, esz, pp=0):
    return ((2 if pp & 1 else 1) * max(GF, GI) * lds_area(ny, W, R) + ny) * esz


# Option bits of a schedule (its last field, the kernels' PP argument):
# 1 double-buffered LDS exchanges, 2 streaming column-pass stores, 4 streaming
# row-pass stores, 8 streaming column-pass loads of P, Q, 16 integrating-factor
# table at radix 16 (FP32), 32 paired rows (two rows as one full-length complex
# transform in the row pass), 64 half-brick row blocks (a row block covers
# by/2 rows of the by-high bricks), 128 P, Q in bricks of 2 by rows (the row
# pass stores and the column pass loads them so).  The row kernel reads bits
# 1, 4, 32, 64 and 128, the column kernel 1, 2, 8, 16 and 128.
ROW_OPTS, COL_OPTS = 485, 411


def row_rows(by, pp):
    """Rows per row block of a schedule (half the brick height with option 64)."""
    return by // 2 if pp & 64 else by

# Schedules that also get a paired-rows variant (options | 32), per precision
# and size: the default and the tuned single-run and batched choices.
PAIRED_BASES = {
    32: {256: [(4, 2, 4, 3, 4, 2, 3, 2, 0), (4, 2, 8, 1, 8, 1, 1, 2, 0)],
         512: [(4, 2, 4, 3, 4, 2, 3, 1, 1), (4, 4, 8, 1, 16, 1, 1, 4, 0)],
         1024: [(4, 2, 4, 3, 4, 2, 1, 2, 15), (4, 2, 8, 1, 8, 1, 1, 2, 0), (4, 2, 8, 3, 8, 2, 3, 2, 0)],
         2048: [(4, 2, 16, 1, 16, 1, 1, 2, 16), (4, 4, 8, 1, 16, 1, 1, 4, 14), (4, 4, 8, 1, 16, 1, 1, 4, 0)],
         4096: [(4, 4, 16, 1, 16, 1, 1, 2, 12)]},
    64: {256: [(4, 2, 4, 3, 4, 2, 3, 2, 0)], 512: [(4, 2, 4, 3, 4, 2, 3, 2, 0)],
         1024: [(4, 2, 4, 3, 4, 2, 3, 2, 0)], 2048: [(4, 2, 8, 1, 8, 1, 1, 1, 0)]},
}
# FP32 schedules that also get the cache-policy variants (options | 2 ... 14,
# and | 16 ... 30 for radix-16 column passes): the fastest single-run and
# batched schedules of each size, and the engine's choices for the mixed
# ensemble (4 x 4 bricks, 4-column blocks).
STORE_VARIANT_BASES = {
    256: [(4, 2, 4, 3, 4, 2, 3, 2, 1), (4, 4, 8, 1, 16, 1, 1, 4, 0)],
    512: [(4, 2, 4, 3, 4, 2, 3, 1, 0), (4, 4, 8, 3, 8, 2, 3, 4, 0), (4, 4, 8, 1, 8, 1, 1, 4, 0),
          (4, 4, 8, 1, 16, 1, 1, 4, 0)],
    1024: [(4, 2, 4, 3, 4, 2, 3, 2, 1), (4, 2, 8, 1, 8, 1, 1, 2, 0), (4, 4, 8, 1, 16, 1, 1, 4, 0)],
    2048: [(4, 2, 16, 1, 16, 1, 1, 2, 0), (4, 2, 16, 1, 16, 1, 1, 2, 1), (4, 4, 8, 1, 16, 1, 1, 4, 0)],
    4096: [(4, 4, 16, 1, 16, 1, 1, 2, 0)],
}
# FP32 schedules that also get half-brick row-block variants (options | 64):
# the row pass with half the rows per block (radix-8 rows of 2 x 2048 points
# in 512 lanes) under the column pass of the tuned 4 x 4 bricks.  On four-row
# bricks they measured slower than whole-brick row blocks at 1024^2 and
# 2048^2, single and batched, so only 4096^2 has them there; the eight-row
# bricks below always use them.
HALF_ROW_BASES = {
    4096: [(4, 4, 8, 1, 16, 1, 1, 2, 8), (4, 4, 8, 1, 16, 1, 1, 2, 12), (4, 4, 16, 1, 16, 1, 1, 2, 12),
           (4, 4, 8, 1, 16, 1, 1, 2, 44), (4, 2, 8, 1, 16, 1, 1, 2, 12)],
}
# Schedules that also get a tall-P, Q variant (options | 128: P, Q in bricks
# of 2 by rows, so the column pass reads them in longer contiguous runs): the
# 4096^2 single-run choice, where it wins (column pass 138 against 147 us);
# for the tuned 256^2-2048^2 single-run, batched and mixed-ensemble choices
# it measured within noise of them, so they are not instantiated.
TALL_PQ_BASES = {
    4096: [(4, 4, 8, 1, 16, 1, 1, 2, 76)],
}
# Per-field layouts (options | 64 | 256 on 8-row bricks): U, V, W (read by the
# row pass) in bricks of the row tile's 4 rows, so a tile reads one contiguous
# run, while P, Q, acc and omega (read by the column pass) sit in the 8-row
# bricks.  The 1024^2 and 2048^2 single runs and the 512^2 class of the mixed
# ensemble take them (1.8 %, 1.8 % and 1.0 % per step); at 256^2, 512^2 single
# and 4096^2 (where the column pass's scattered U, V, W stores cost 50 us) they
# lost.
SPLIT_LAYOUT_BASES = {
    512: [(8, 4, 8, 1, 8, 1, 1, 4, 322)],
    1024: [(8, 2, 4, 3, 4, 2, 1, 2, 335), (8, 2, 4, 3, 4, 2, 1, 2, 323)],
    2048: [(8, 2, 16, 1, 16, 1, 1, 2, 336), (8, 2, 16, 1, 16, 1, 1, 2, 320)],
}
# Eight-row bricks (128- or 256-byte bricks for the column pass) with
# half-brick (four-row) row tiles: the engine's 2048^2 choice (2.5-3.3 %
# faster than 4 x 4 bricks for batches of 4 and 8) and its neighbours; at
# 256^2-1024^2 (single runs and the mixed ensemble) they tied or lost.
TALL_BRICK_BASES = {
    1024: [(8, 2, 4, 3, 4, 2, 1, 2, 15)],
    2048: [(8, 2, 16, 1, 16, 1, 1, 2, 16), (8, 4, 8, 1, 16, 1, 1, 4, 14), (8, 4, 8, 1, 16, 1, 1, 4, 10),
           (8, 4, 8, 1, 16, 1, 1, 4, 46)],
}


# Mixes instantiated as one launch per pass (the ensemble's single-launch
# mode): per precision, lists of (n, schedule) for n x n classes, in launch
# order (largest first).  Every schedule of a mix has 256-lane blocks in both
# passes: one launch has one block size.
MULTI = {
    32: [[(1024, (4, 2, 8, 1, 8, 1, 1, 2, 0)), (512, (4, 4, 4, 1, 8, 1, 1, 4, 0)), (256, (8, 4, 4, 1, 4, 1, 1, 4, 0))]],
}


def default_schedule(prec, nx, ny):
    """(by, bx, row_r, row_g, col_r, col_gf, col_gi, col_w, pp)"""
    nx2 = nx // 2
    if prec == 32:
        by = 2 if nx2 >= 2048 else 4
        bx = 2
        row = (4, 3) if nx2 == 16 else ((8, 1) if nx2 >= 2048 else (8, 3))
        col = (16, 1, 1) if ny >= 4096 else (8, 2, 3)
    else:
        by = 2 if nx2 >= 2048 else 4
        bx = 1 if ny >= 4096 else 2
        row = (4, 3) if nx2 == 16 else ((8, 1) if nx2 >= 1024 else (8, 3))
        col = (8, 1, 1) if ny >= 2048 else (8, 2, 3)
    return (by, bx) + row + col + (bx, 0)


def candidates(prec, n):
    """Alternative schedules for the autotuner (square n x n only)."""
    out = [default_schedule(prec, n, n)]
    nx2 = n // 2
    if prec == 32:
        alts = [(4, 2, 4, 3, 4, 2, 3), (4, 4, 8, 3, 8, 2, 3), (2, 4, 8, 3, 8, 2, 3), (4, 2, 8, 1, 8, 1, 1),
                (4, 2, 8, 3, 16, 1, 1), (2, 2, 8, 3, 8, 2, 3), (4, 2, 4, 3, 8, 2, 3), (4, 2, 8, 3, 4, 2, 3)]
        narrow = [(4, 2, 4, 3, 4, 2, 3, 1), (4, 2, 8, 1, 8, 1, 1, 1), (4, 2, 8, 3, 8, 2, 3, 1),
                  (4, 4, 8, 3, 8, 2, 3, 2), (2, 2, 8, 1, 16, 1, 1, 1), (4, 2, 4, 3, 8, 2, 3, 1),
                  (4, 2, 4, 3, 16, 1, 1, 1), (4, 2, 8, 3, 16, 1, 1, 1), (4, 2, 4, 3, 8, 1, 1, 1),
                  (2, 2, 8, 1, 8, 1, 1, 2), (2, 2, 8, 1, 8, 1, 1, 1), (2, 2, 4, 3, 8, 1, 1, 2),
                  (2, 2, 4, 1, 4, 1, 1, 2), (4, 4, 8, 1, 16, 1, 1, 4), (4, 4, 8, 1, 8, 1, 1, 4),
                  (4, 4, 8, 3, 16, 1, 1, 4), (4, 2, 16, 1, 8, 1, 1, 2), (4, 2, 16, 3, 8, 1, 1, 2),
                  (4, 2, 16, 1, 4, 2, 3, 2)]
    else:
        alts = [(4, 2, 4, 3, 4, 2, 3), (4, 1, 8, 3, 8, 2, 3), (2, 2, 8, 3, 8, 2, 3), (4, 2, 8, 1, 8, 1, 1),
                (4, 2, 4, 1, 4, 1, 1), (2, 2, 4, 3, 8, 1, 1)]
        narrow = [(4, 2, 4, 3, 4, 2, 3, 1), (4, 2, 8, 3, 8, 2, 3, 1), (4, 2, 8, 1, 8, 1, 1, 1)]
    alts = [a + (a[1], 0) for a in alts] + [a + (0,) for a in narrow]
    if prec == 32:   # radix-16 rows and columns with 4-row bricks (2048^2, 4096^2)
        alts += [(4, 2, 16, 1, 16, 1, 1, 2, 0), (4, 2, 16, 1, 16, 1, 1, 1, 0), (2, 2, 16, 1, 16, 1, 1, 2, 0),
                 (4, 4, 16, 1, 16, 1, 1, 2, 0), (4, 4, 16, 1, 16, 1, 1, 4, 0)]
    # double-buffered LDS exchanges for the single-run schedules of the small sizes
    # (one block per CU there, so the doubled LDS costs no occupancy)
    if prec == 32:
        alts += [(4, 2, 4, 3, 4, 2, 3, 2, 1), (4, 2, 4, 3, 4, 2, 3, 1, 1), (4, 2, 8, 3, 8, 2, 3, 2, 1),
                 (4, 2, 16, 1, 8, 1, 1, 2, 1), (4, 2, 8, 1, 8, 1, 1, 2, 1)]
    else:
        alts += [(4, 2, 4, 3, 4, 2, 3, 2, 1), (4, 1, 8, 3, 8, 2, 3, 1, 1), (4, 2, 8, 1, 8, 1, 1, 1, 1)]
    if prec == 32 and n in (512, 1024):
        # one field per exchange on one side (stores of a finished field drain
        # while the next one is transformed), with the cache options of the
        # single-run winner
        for rg, gf, gi in ((1, 2, 3), (3, 2, 1), (3, 1, 1), (1, 2, 1), (1, 1, 1)):
            for o in (1, 3, 7, 11, 15):
                alts.append((4, 2, 4, rg, 4, gf, gi, 2 if n == 1024 else 1, o))
    for base in PAIRED_BASES.get(prec, {}).get(n, []):
        alts.append(base[:8] + (base[8] | 32,))
    for mix in MULTI.get(prec, []):   # the schedules of the mixed launches
        alts += [s for m, s in mix if m == n]
    if prec == 32:
        for b in HALF_ROW_BASES.get(n, []) + TALL_BRICK_BASES.get(n, []):
            alts.append(b[:8] + (b[8] | 64,))
        for b in TALL_PQ_BASES.get(n, []):
            alts.append(b[:8] + (b[8] | 128,))
        alts += SPLIT_LAYOUT_BASES.get(n, [])
        for b in STORE_VARIANT_BASES.get(n, []):
            opts = list(range(2, 16, 2)) + ([16 + o for o in range(0, 16, 2)] if b[4] == 16 else [])
            alts += [b[:8] + (b[8] | o,) for o in opts]
    for s in alts:
        if s not in out and feasible(prec, n, n, s):
            out.append(s)
    return [s for s in out if feasible(prec, n, n, s)]


def feasible(prec, nx, ny, s):
    by, bx, rr, rg, cr, gf, gi, cw, pp = s
    esz = 8 if prec == 32 else 16
    nx2 = nx // 2
    if rr > nx2 or cr > ny:
        return False
    tr, tc = nx2 // rr, ny // cr
    if tr % bx or nx2 % bx or ny % by or bx % cw or tc % by:
        return False
    if pp & 64 and by % 2:   # half-brick row blocks: an even brick height
        return False
    if pp & 128 and (pp & 32 or tc % (2 * by) or ny % (2 * by)):   # tall P, Q bricks
        return False
    if pp & 256 and (pp & 32 or not pp & 64):   # U, V, W in half-height bricks: half-brick tiles
        return False
    rows = row_rows(by, pp)
    if pp & 32 and (rows % 2 or rr < 2):   # paired rows: pairs of block rows, mirrored half of each lane
        return False
    if rows * tr > 512 or cw * tc > 512:
        return False
    if lds_row(nx2, rows, rr, rg, esz, pp) > LDS_MAX or lds_col(ny, cw, cr, gf, gi, esz, pp) > LDS_MAX:
        return False
    return True


HDR = """// {what} schedule and launch tables (the kernels are instantiated in
// kernels_{tag}_k*.hip, compiled in parallel).
// Generated by gen_kernel_tables.py -- edit that script, not this file.
// Row kernel:    (Nx/2, RB = rows per block (the brick height, half of it with PP 64), R, BX = brick width, G, PP)
// Column kernel: (Ny, W = columns per block, R, BY, BX = brick height/width, GF, GI, PP)
// Schedule:      (Nx/2, Ny, BY, BX, row R, row G, col R, col GF, col GI, col W, PP, default)
// PP: option bits (1 double-buffered LDS, 2 / 4 streaming column / row stores,
// 8 streaming column-pass P, Q loads, 16 integrating-factor table at radix 16,
// 32 paired rows, 64 half-brick row blocks, 128 P, Q in bricks twice as tall)
#include "kernels_inst.hpp"

namespace af {{
{externs}
namespace {{
using T = {T};
const RowCfg kRow[] = {{
{rows}
}};
const ColCfg kCol[] = {{
{cols}
}};
const Schedule kSched[] = {{
{scheds}
}};
}}  // namespace

const RowCfg* row_cfgs_{tag}(int* n) {{ *n = (int)(sizeof(kRow) / sizeof(kRow[0])); return kRow; }}
const ColCfg* col_cfgs_{tag}(int* n) {{ *n = (int)(sizeof(kCol) / sizeof(kCol[0])); return kCol; }}
const Schedule* schedules_{tag}(int* n) {{ *n = (int)(sizeof(kSched) / sizeof(kSched[0])); return kSched; }}
ConvLaunch to_bricks_{tag}() {{ return launch_to_bricks<T>; }}
ConvLaunch from_bricks_{tag}() {{ return launch_from_bricks<T>; }}
{ifac}
}}  // namespace af
"""


def gen(prec, outdir):
    tag = "f32" if prec == 32 else "f64"
    T = "float" if prec == 32 else "double"
    scheds = []
    for nx in SIZES:
        for ny in SIZES:
            d = default_schedule(prec, nx, ny)
            assert feasible(prec, nx, ny, d), (prec, nx, ny, d)
            scheds.append((nx // 2, ny) + d + (1,))
            if nx == ny and nx >= 256:
                for s in candidates(prec, nx)[1:]:
                    scheds.append((nx // 2, ny) + s + (0,))
    rows, cols = set(), set()
    for (nx2, ny, by, bx, rr, rg, cr, gf, gi, cw, pp, _) in scheds:
        rows.add((nx2, row_rows(by, pp), rr, bx, rg, pp & ROW_OPTS))
        cols.add((ny, cw, cr, by, bx, gf, gi, pp & COL_OPTS))
    rl = [f"    {{{a}, {b}, {r}, {c}, {b * (a // r)}, launch_row<T, {a}, {b}, {r}, {c}, {g}, {pp}>, {g}, {pp}}},"
          for (a, b, r, c, g, pp) in sorted(rows)]
    cl = [f"    {{{a}, {w}, {r}, {by}, {bx}, {w * (a // r)}, launch_col<T, {a}, {w}, {r}, {by}, {bx}, {gf}, {gi}, {pp}>, "
          f"{gf}, {gi}, {pp}}}," for (a, w, r, by, bx, gf, gi, pp) in sorted(cols)]
    sl = ["    {" + ", ".join(map(str, t)) + "}," for t in scheds]
    ifac = "IfacLaunch ifac_init_f32() { return launch_ifac_init; }\n" if prec == 32 else ""
    # explicit instantiations, spread over parts of similar weight (a column
    # launcher instantiates 5 stage kernels, a row launcher 1)
    row_sig = "(const void*, int, int, const void*, hipStream_t)"
    col_sig = "(const void*, int, int, int, const void*, hipStream_t)"
    insts = [(1, f"hipError_t launch_row<{T}, {a}, {b}, {r}, {c}, {g}, {pp}>{row_sig}")
             for (a, b, r, c, g, pp) in sorted(rows)]
    insts += [(5, f"hipError_t launch_col<{T}, {a}, {w}, {r}, {by}, {bx}, {gf}, {gi}, {pp}>{col_sig}")
              for (a, w, r, by, bx, gf, gi, pp) in sorted(cols)]
    nparts = PARTS[prec]
    parts = [[] for _ in range(nparts)]
    load = [0] * nparts
    for wgt, sig in sorted(insts, key=lambda x: -x[0]):
        k = load.index(min(load))
        parts[k].append(sig)
        load[k] += wgt
    externs = "\n".join(f"extern template {sig};" for _, sig in insts)
    text = HDR.format(what=f"FP{prec}", T=T, tag=tag, rows="\n".join(rl), cols="\n".join(cl), scheds="\n".join(sl),
                      ifac=ifac, externs=externs)
    with open(os.path.join(outdir, f"kernels_{tag}.hip"), "w") as f:
        f.write(text)
    for k, sigs in enumerate(parts):
        body = "\n".join(f"template {sig};" for sig in sigs)
        with open(os.path.join(outdir, f"kernels_{tag}_k{k}.hip"), "w") as f:
            f.write(PART.format(what=f"FP{prec}", k=k, n=nparts, body=body))
    return len(rl), len(cl), len(sl)


PARTS = {32: 8, 64: 4}
PART = """// {what} stage-kernel instantiations, part {k} of {n}.
// Generated by gen_kernel_tables.py -- edit that script, not this file.
#include "kernels_inst.hpp"

namespace af {{
{body}
}}  // namespace af
"""


MULTI_FILE = """// Mixed launches: one launch per pass over the size classes of each mix.
// Generated by gen_kernel_tables.py -- edit that script, not this file.
#include "multi_pass.hpp"

namespace af {{
namespace {{
const MultiCfg kMulti[] = {{
{entries}
}};
}}  // namespace

const MultiCfg* multi_cfgs(int* n) {{
  *n = {count};
  return {ptr};
}}

}}  // namespace af
"""


def gen_multi(outdir):
    entries = []
    for prec, mixes in MULTI.items():
        T = "float" if prec == 32 else "double"
        for mix in mixes:
            assert len(mix) <= 4
            cls, rows, cols = [], [], []
            for n, s in mix:
                by, bx, rr, rg, cr, gf, gi, cw, pp = s
                assert feasible(prec, n, n, s) and s in candidates(prec, n), (n, s)
                assert by * (n // 2 // rr) == 256 and cw * (n // cr) == 256, (n, s)
                assert not pp & 480, "mixed launches call the unpaired whole-brick row body"
                cls.append("{" + ", ".join(map(str, (n // 2, n, by, bx, rr, rg, cr, gf, gi, cw, pp, 0))) + "}")
                rows.append(f"RowPass<{T}, {n // 2}, {by}, {rr}, {bx}, {rg}, {pp & ROW_OPTS}>")
                cols.append(f"ColCfgT<{T}, {n}, {cw}, {cr}, {by}, {bx}, {gf}, {gi}, {pp & COL_OPTS}>")
            entries.append(f"    {{{prec}, {len(mix)}, {{{', '.join(cls)}}},\n     launch_row_multi<{', '.join(rows)}>,\n"
                           f"     launch_col_multi<{', '.join(cols)}>}},")
    text = MULTI_FILE.format(entries="\n".join(entries) if entries else "    {}",
                             count=len(entries), ptr="kMulti" if entries else "nullptr")
    with open(os.path.join(outdir, "kernels_multi.hip"), "w") as f:
        f.write(text)
    return len(entries)


if __name__ == "__main__":
    if len(sys.argv) != 2:
        sys.exit("usage: gen_kernel_tables.py OUTDIR")
    os.makedirs(sys.argv[1], exist_ok=True)
    for prec in (32, 64):
        print(prec, gen(prec, sys.argv[1]))
    print("mixes", gen_multi(sys.argv[1]))
