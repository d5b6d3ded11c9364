"""Writes the golden fixtures under tests/golden/ with the reference's own CPU tools.

Run in the build container only (it needs the compiled reference tools from
`make -C oracle ref`, which read /root/reference):

    python tests/golden/make_fixtures.py

Each case is one `<case>.npz` holding
  * `src`, `dst`   (uint32): the compressor's input edge list, sorted by (src, dst).
                   It is also the golden CSR: decoding the files must give back exactly this.
  * `n`            (int64): max id + 1, the vertex count the compressor derives
                   (LE/chunk_compress/compressor.hpp:106-107).
  * `le_graph`, `le_offset`, `gt_graph`, `gt_offset` (uint8): the files written by the
                   reference LE and GT compressors (LE/chunk_compress/compressor.hpp:130-154,
                   GT/chunk_compress/compressor.hpp:120-144), "normal" encode type.
graph_sort cases (`gs_<case>.npz`) hold the raw input (`raw_u`, `raw_v`, int64), the mode
(`directed`) and the reference graph_sort output (`out_u`, `out_v`)
(LE/chunk_compress/graph_sort.cpp:22-82).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")


def _tool(name):
    p = os.path.join(REF_BIN, name)
    if not os.path.exists(p):
        sys.exit(f"missing {p}: run `make -C oracle ref` first")
    return p


def _compress(src, dst, tmp, tag):
    txt = os.path.join(tmp, tag + ".txt")
    with open(txt, "w") as f:
        for u, v in zip(src.tolist(), dst.tolist()):
            f.write(f"{u} {v}\n")
    out = {}
    for fmt in ("le", "gt"):
        prefix = os.path.join(tmp, f"{tag}_{fmt}")
        subprocess.run([_tool(f"{fmt}_compressor"), txt, prefix], check=True,
                       stdout=subprocess.DEVNULL)
        out[f"{fmt}_graph"] = np.fromfile(prefix + ".graph", dtype=np.uint8)
        out[f"{fmt}_offset"] = np.fromfile(prefix + ".offset", dtype=np.uint8)
    return out


def _save_case(name, adj, tmp):
    """adj: dict vertex -> sorted list of distinct neighbours (no self loops)."""
    src, dst = [], []
    for u in sorted(adj):
        for v in adj[u]:
            src.append(u)
            dst.append(v)
    src = np.asarray(src, dtype=np.uint32)
    dst = np.asarray(dst, dtype=np.uint32)
    n = int(max(src.max(), dst.max())) + 1
    files = _compress(src, dst, tmp, name)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), src=src, dst=dst,
                        n=np.int64(n), **files)
    print(f"{name}: n={n} m={len(src)} le={len(files['le_graph'])}B "
          f"gt={len(files['gt_graph'])}B ubl_le={files['le_offset'][0]} ubl_gt={files['gt_offset'][0]}")


def _graph_sort(raw_u, raw_v, directed, tmp, tag):
    txt = os.path.join(tmp, tag + ".raw")
    with open(txt, "w") as f:
        for u, v in zip(raw_u.tolist(), raw_v.tolist()):
            f.write(f"{u} {v}\n")
    out = os.path.join(tmp, tag + ".sorted")
    subprocess.run([_tool("graph_sort"), txt, out, "d" if directed else "u"], check=True,
                   stdout=subprocess.DEVNULL)
    a = np.loadtxt(out, dtype=np.int64, ndmin=2)
    return a[:, 0], a[:, 1]


def _rmat(scale, ef, rng, a=0.57, b=0.19, c=0.19):
    m = ef << scale
    u = np.zeros(m, dtype=np.int64)
    v = np.zeros(m, dtype=np.int64)
    for _ in range(scale):
        r = rng.random(m)
        bu = (r >= a + b).astype(np.int64)
        bv = (((r >= a) & (r < a + b)) | (r >= a + b + c)).astype(np.int64)
        u = (u << 1) | bu
        v = (v << 1) | bv
    return u, v


def main():
    rng = np.random.default_rng(20240611)
    with tempfile.TemporaryDirectory() as tmp:
        # 1. The hand-derived example: v = 5, N = [1, 3, 8, 9, 20].
        _save_case("example", {5: [1, 3, 8, 9, 20]}, tmp)

        # 2. Degree sweep: d in {0..11} plus k^2-1, k^2, k^2+1 for several k, with the
        #    neighbour set all below v, all above v, straddling v, or a run of consecutive
        #    ids (zero gaps, zero residuals).
        nv = 3000
        degs = list(range(0, 12))
        for k in (4, 5, 6, 7, 8, 10, 11, 12, 16, 23):
            degs += [k * k - 1, k * k, k * k + 1]
        adj = {}
        modes = ("below", "above", "straddle", "run", "random")
        vid = 1
        for d in degs:
            for mode in modes:
                v = vid * 7 + 1000
                vid += 1
                if d == 0:
                    adj[v] = []
                    continue
                if mode == "below":
                    cand = np.arange(0, v)
                elif mode == "above":
                    cand = np.arange(v + 1, nv)
                elif mode == "straddle":
                    cand = np.concatenate([np.arange(v - 3 * d, v), np.arange(v + 1, v + 3 * d + 1)])
                    cand = cand[(cand >= 0) & (cand < nv)]
                elif mode == "run":
                    start = v + 1 if (d % 2) else max(0, v - d - 5)
                    cand = np.arange(start, start + d + 1)
                    cand = cand[cand != v][:d]
                    adj[v] = sorted(int(x) for x in cand)
                    continue
                else:
                    cand = np.concatenate([np.arange(0, v), np.arange(v + 1, nv)])
                pick = rng.choice(cand, size=d, replace=False)
                adj[v] = sorted(int(x) for x in pick)
        # vertices of degree 0 in the middle and a trailing vertex that is only a target
        adj[2] = [nv + 5]
        _save_case("degrees", {k: v for k, v in adj.items() if v}, tmp)

        # 3. A hub large enough that the longest record needs > 16 bits (ubl > 16).
        nh = 70000
        hub = 31337
        adj = {hub: sorted(int(x) for x in rng.choice(np.setdiff1d(np.arange(nh), [hub]),
                                                       size=30000, replace=False))}
        for v in rng.choice(nh, size=400, replace=False).tolist():
            if v == hub:
                continue
            d = int(rng.integers(1, 40))
            cand = rng.choice(nh, size=d + 1, replace=False)
            adj[v] = sorted(int(x) for x in cand if x != v)[:d]
        _save_case("hub", adj, tmp)

        # 4. A record-length field narrower than 8 bits with padding >= ubl: the reader's
        #    n = floor(8 (bytes - 1) / ubl) (LE/src/grass.cu:1088) overcounts by 2 here.
        _save_case("overcount", {0: [2]}, tmp)

        # 5. graph_sort (undirected) over a small random multigraph with self loops and
        #    duplicates, n < 256, then both compressors on its output.
        raw_u = rng.integers(0, 80, size=400)
        raw_v = rng.integers(0, 80, size=400)
        raw_u[:7] = raw_v[:7]  # self loops, including on ids that appear nowhere else
        raw_u[7] = 1000
        raw_v[7] = 1000
        out_u, out_v = _graph_sort(raw_u, raw_v, False, tmp, "small")
        np.savez_compressed(os.path.join(HERE, "gs_small_u.npz"), raw_u=raw_u, raw_v=raw_v,
                            directed=np.bool_(False), out_u=out_u, out_v=out_v)
        adj = {}
        for u, v in zip(out_u.tolist(), out_v.tolist()):
            adj.setdefault(u, []).append(v)
        _save_case("small_u", adj, tmp)

        # 6. graph_sort, directed mode, same raw input.
        out_u, out_v = _graph_sort(raw_u, raw_v, True, tmp, "small_d")
        np.savez_compressed(os.path.join(HERE, "gs_small_d.npz"), raw_u=raw_u, raw_v=raw_v,
                            directed=np.bool_(True), out_u=out_u, out_v=out_v)

        # 7. R-MAT scale 12, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05): graph_sort 'u',
        #    then both compressors (skewed degrees, hubs, many d = 1 and d = 2 records).
        ru, rv = _rmat(12, 16, rng)
        out_u, out_v = _graph_sort(ru, rv, False, tmp, "rmat12")
        np.savez_compressed(os.path.join(HERE, "gs_rmat12_u.npz"), raw_u=ru, raw_v=rv,
                            directed=np.bool_(False), out_u=out_u, out_v=out_v)
        adj = {}
        for u, v in zip(out_u.tolist(), out_v.tolist()):
            adj.setdefault(u, []).append(v)
        _save_case("rmat12", adj, tmp)


if __name__ == "__main__":
    main()
