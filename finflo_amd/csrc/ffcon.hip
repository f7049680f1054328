// ffcon.hip — one-to-one block connections on the device: the halo exchange of BOUNDI
// (EXPTOT/IMPVO2 gluebm.f:6819-7262, CONNEC ns3c.f:8605-8857) as a gather of precomputed
// (ghost cell, source cell) offset pairs, and the multi-block cycle (CYCLES' block loop with
// BOUNDI(1) once per cycle, ns3c.f:556, 1400-2195).
// Included at the end of ffhip.hip (it uses the context and launch helpers defined there).

constexpr int HALO_MAXF = 24;
struct HaloFields {
    int n;
    int f[HALO_MAXF];
};

// the fields the flow exchange copies (the solved variables, what STATIM derives from them,
// and the eddy-viscosity fields of TURBDI) and those of the geometry exchange
static const int kHaloFlow[] = {FF_PDIFF, FF_TEMP, FF_U, FF_V, FF_W, FF_RK, FF_REPS,
                                FF_P, FF_RO, FF_E, FF_C, FF_VIS, FF_CH, FF_CP, FF_DRDP,
                                FF_DRDH, FF_RM, FF_RN, FF_RW, FF_EPS2, FF_VIST};
static const int kHaloGeom[] = {FF_VOL, FF_XC, FF_YC, FF_ZC, FF_DISTW, FF_D1, FF_D2, FF_D3};

// thread per (ghost, source) pair: every listed field of the ghost cell of `dst` takes the
// value of the source cell of `src`
__global__ void __launch_bounds__(256) k_halo(Blk dst, Blk src, const int *map, int n, HaloFields hf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int d = map[2 * t], s = map[2 * t + 1];
#pragma unroll 4
    for (int q = 0; q < hf.n; q++) FLD(dst, hf.f[q])[d] = FLD(src, hf.f[q])[s];
}

// The same copies for every connection of a level in one launch: workgroup w takes pairs
// wg[w].y .. wg[w].y + 255 of connection wg[w].x.  Every ghost cell has one source, and the
// sources are interior cells, which no pair writes, so the order of the copies does not
// matter and the result is that of the launches one connection after the other.
__global__ void __launch_bounds__(256) k_halo_batch(const Blk *blks, const HaloConnDev *conns,
                                                    const int2 *wg, HaloFields hf)
{
    const int2 w = wg[blockIdx.x];
    const HaloConnDev cn = conns[w.x];
    const int t = w.y + threadIdx.x;
    if (t >= cn.n) return;
    const int d = cn.map[2 * t], s = cn.map[2 * t + 1];
    const Blk &dst = blks[cn.dst], &src = blks[cn.src];
#pragma unroll 4
    for (int q = 0; q < hf.n; q++) FLD(dst, hf.f[q])[d] = FLD(src, hf.f[q])[s];
}

// the in-plane extents of a face: (J, K) on an I-face, (I, K) on a J-face, (I, J) on a K-face
static void face_extent(const Blk &b, int face, int *ea, int *eb)
{
    const int dir = (face - 1) % 3;
    *ea = dir == 0 ? b.jmax : b.imax;
    *eb = dir == 2 ? b.jmax : b.kmax;
}

// 0-based offset of the cell at in-plane (a, b) on `face`, `depth` cells in from the face
// (1, 2: the interior layers; 0, -1: the ghost layers behind it)
static int face_cell(const Blk &b, int face, int a, int bb, int depth)
{
    const int dir = (face - 1) % 3;
    const int nd = dir == 0 ? b.imax : (dir == 1 ? b.jmax : b.kmax);
    const int nrm = face <= 3 ? depth : nd + 1 - depth;
    int i, j, k;
    if (dir == 0) { i = nrm; j = a; k = bb; }
    else if (dir == 1) { i = a; j = nrm; k = bb; }
    else { i = a; j = bb; k = nrm; }
    return (k + 1) * b.il + (j + 1) * b.is + (i + 1);
}

extern "C" int ff_con_register(ff_ctx *c, int level, int block, const ff_con_patch *p, int n)
{
    if (c) c->gen++;
    if (!c || n < 0 || (n > 0 && !p)) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    const Blk &A = r->dev;
    std::vector<int> pairs;
    std::vector<size_t> off;
    for (int l = 0; l < n; l++) {
        const ff_con_patch &h = p[l];
        const std::string ls = std::to_string(l);
        if (h.iface < 1 || h.iface > 6 || h.nface < 1 || h.nface > 6)
            return fail(c, FF_E_ARG, "ff_con_register: patch %s: face outside 1..6", ls.c_str());
        BlockRec *nr = find(c, level, h.nblock);
        if (!nr) return fail(c, FF_E_STATE, "ff_con_register: patch %s: connected block not registered", ls.c_str());
        const Blk &B = nr->dev;
        int ea, eb, fa, fb;
        face_extent(A, h.iface, &ea, &eb);
        face_extent(B, h.nface, &fa, &fb);
        if (h.ka1 < 1 || h.ka2 > ea || h.ka1 > h.ka2 || h.kb1 < 1 || h.kb2 > eb || h.kb1 > h.kb2)
            return fail(c, FF_E_ARG, "ff_con_register: patch %s outside its face", ls.c_str());
        if ((h.da != 1 && h.da != -1) || (h.db != 1 && h.db != -1) || (h.swap != 0 && h.swap != 1))
            return fail(c, FF_E_ARG, "ff_con_register: patch %s: orientation (da, db = +-1, swap = 0/1)", ls.c_str());
        const int ndA = (h.iface - 1) % 3 == 0 ? A.imax : ((h.iface - 1) % 3 == 1 ? A.jmax : A.kmax);
        const int ndB = (h.nface - 1) % 3 == 0 ? B.imax : ((h.nface - 1) % 3 == 1 ? B.jmax : B.kmax);
        if (ndA < 2 || ndB < 2)
            return fail(c, FF_E_UNSUP, "ff_con_register: patch %s: blocks must be two cells thick across a connection", ls.c_str());
        const int np = h.ka2 - h.ka1 + 1, nq = h.kb2 - h.kb1 + 1;
        // the far corner must lie on the connected face too
        for (int corner = 0; corner < 4; corner++) {
            const int pp = corner & 1 ? np - 1 : 0, qq = corner & 2 ? nq - 1 : 0;
            const int na = h.swap ? h.na0 + h.da * qq : h.na0 + h.da * pp;
            const int nb = h.swap ? h.nb0 + h.db * pp : h.nb0 + h.db * qq;
            if (na < 1 || na > fa || nb < 1 || nb > fb)
                return fail(c, FF_E_ARG, "ff_con_register: patch %s maps outside the connected face", ls.c_str());
        }
        off.push_back(pairs.size() / 2);
        for (int layer = 1; layer <= 2; layer++)
            for (int q = 0; q < nq; q++)
                for (int pp = 0; pp < np; pp++) {
                    const int na = h.swap ? h.na0 + h.da * q : h.na0 + h.da * pp;
                    const int nb = h.swap ? h.nb0 + h.db * pp : h.nb0 + h.db * q;
                    pairs.push_back(face_cell(A, h.iface, h.ka1 + pp, h.kb1 + q, 1 - layer));
                    pairs.push_back(face_cell(B, h.nface, na, nb, layer));
                }
    }
    off.push_back(pairs.size() / 2);
    // the sources are interior layers 1 and 2 (two cells thick, checked above); a ghost cell
    // named by two patches (overlapping ranges on one face) makes the copies order-dependent
    std::vector<int> dsts(pairs.size() / 2);
    for (size_t q = 0; q < dsts.size(); q++) dsts[q] = pairs[2 * q];
    std::sort(dsts.begin(), dsts.end());
    const bool unique = std::adjacent_find(dsts.begin(), dsts.end()) == dsts.end();
    int *map = nullptr;
    if (!pairs.empty()) {
        if (hipMalloc(&map, sizeof(int) * pairs.size()) != hipSuccess)
            return fail(c, FF_E_NOMEM, "connection maps%s");
        if (hipMemcpy(map, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(map);
            return fail(c, FF_E_HIP, "connection maps: copy failed%s");
        }
    }
    if (r->con_map) (void)hipFree(r->con_map);
    r->con_map = map;
    r->cons.assign(p, p + n);
    r->con_off = off;
    r->con_unique = unique;
    return FF_OK;
}

// The device tables of the level's batched exchange, (re)built when the registrations have
// changed since the last build.  Returns 1 when the batch can run, 0 when the exchange must go
// connection by connection (a ghost cell named twice, or a stale table while a graph is being
// captured, when nothing may be allocated), or an error.
static int halo_batch_ready(ff_ctx *c, int level)
{
    HaloBatch &hb = c->halo_batch[level];
    if (hb.gen == c->gen) return hb.nwg > 0 ? 1 : 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(c, hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return 0;
    std::vector<Blk> blks;
    std::map<int, int> index;          // block number -> position in blks
    // the context's blocks are keyed by (level, block number), block number in the low word
    auto number = [](long long k) { return (int)(unsigned)(k & 0xffffffffLL); };
    for (auto &kv : c->blocks)
        if (kv.second.level == level) {
            index[number(kv.first)] = (int)blks.size();
            blks.push_back(kv.second.dev);
        }
    std::vector<HaloConnDev> conns;
    std::vector<int2> wg;
    bool ok = true;
    for (auto &kv : c->blocks) {
        const BlockRec &r = kv.second;
        if (r.level != level || r.cons.empty()) continue;
        ok = ok && r.con_unique;
        for (size_t l = 0; l < r.cons.size(); l++) {
            auto it = index.find(r.cons[l].nblock);
            if (it == index.end())
                return fail(c, FF_E_STATE, "ff_halo_exchange: connected block no longer registered%s");
            const int np = (int)(r.con_off[l + 1] - r.con_off[l]);
            if (np == 0) continue;
            const int ci = (int)conns.size();
            conns.push_back({index[number(kv.first)], it->second, r.con_map + 2 * r.con_off[l], np});
            for (int first = 0; first < np; first += 256) wg.push_back(make_int2(ci, first));
        }
    }
    if (hb.blks) {
        // the last exchange that read the old tables ran on this stream
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(hb.blks);
    }
    if (hb.conns) (void)hipFree(hb.conns);
    if (hb.wg) (void)hipFree(hb.wg);
    hb = HaloBatch{};
    if (!ok || wg.empty()) {
        hb.gen = c->gen;
        return 0;
    }
    if (hipMalloc(&hb.blks, sizeof(Blk) * blks.size()) != hipSuccess ||
        hipMalloc(&hb.conns, sizeof(HaloConnDev) * conns.size()) != hipSuccess ||
        hipMalloc(&hb.wg, sizeof(int2) * wg.size()) != hipSuccess)
        return fail(c, FF_E_NOMEM, "halo exchange tables%s");
    HIPCHK(c, hipMemcpy(hb.blks, blks.data(), sizeof(Blk) * blks.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(hb.conns, conns.data(), sizeof(HaloConnDev) * conns.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(hb.wg, wg.data(), sizeof(int2) * wg.size(), hipMemcpyHostToDevice));
    hb.nwg = (int)wg.size();
    hb.gen = c->gen;
    return 1;
}

static int halo_impl(ff_ctx *c, int level, int what)
{
    HaloFields hf{};
    const int *src = what == 0 ? kHaloFlow : kHaloGeom;
    hf.n = what == 0 ? (int)(sizeof(kHaloFlow) / sizeof(int)) : (int)(sizeof(kHaloGeom) / sizeof(int));
    for (int q = 0; q < hf.n; q++) hf.f[q] = src[q];
    c->lev = level;
    const int batch = halo_batch_ready(c, level);
    if (batch < 0) return batch;
    if (batch == 1) {
        const HaloBatch &hb = c->halo_batch[level];
        return launch(c, "k_halo", k_halo_batch, dim3((unsigned)hb.nwg), dim3(256),
                      (const Blk *)hb.blks, (const HaloConnDev *)hb.conns, (const int2 *)hb.wg, hf);
    }
    for (auto &kv : c->blocks) {
        BlockRec &r = kv.second;
        if (r.level != level || r.cons.empty()) continue;
        for (size_t l = 0; l < r.cons.size(); l++) {
            BlockRec *nr = find(c, level, r.cons[l].nblock);
            if (!nr) return fail(c, FF_E_STATE, "ff_halo_exchange: connected block no longer registered%s");
            const int np = (int)(r.con_off[l + 1] - r.con_off[l]);
            int rc = launch(c, "k_halo", k_halo, dim3((np + 255) / 256), dim3(256), r.dev, nr->dev,
                            (const int *)(r.con_map + 2 * r.con_off[l]), np, hf);
            if (rc) return rc;
        }
    }
    return FF_OK;
}

extern "C" int ff_halo_exchange(ff_ctx *c, int level, int what)
{
    if (!c || (what != 0 && what != 1)) return FF_E_ARG;
    return halo_impl(c, level, what);
}

// The pool streams and fork/join events (created once, default priority).
static int pool_ready(ff_ctx *c)
{
    for (int q = 0; q < ff_ctx::NPOOL; q++)
        if (!c->pool[q] && hipStreamCreateWithFlags(&c->pool[q], hipStreamNonBlocking) != hipSuccess)
            return fail(c, FF_E_HIP, "stream pool%s");
    for (int q = 0; q <= ff_ctx::NPOOL; q++)
        if (!c->pool_ev[q] && hipEventCreateWithFlags(&c->pool_ev[q], hipEventDisableTiming) != hipSuccess)
            return fail(c, FF_E_HIP, "stream pool events%s");
    return FF_OK;
}

// FF_OPT_BLOCK_STREAMS: every cycle forks the pool from the context's stream, runs every
// block's V-cycle on a pool stream (the RESL2 sums of the last cycle into the block's own
// slot, partials in its stream's buffer, no host wait), joins the pool back and exchanges
// the level-1 halos on the context's stream.  Block n goes to pool stream n mod NPOOL.
// (Measured on the 24- and 64-block cases: dealing the blocks out largest first to the
// least loaded stream, or over fewer streams, was slower.)  The staged LU-SGS records are
// allocated before the first fork, so no allocation runs while the pool works.
static int vcycle_blocks_streams(ff_ctx *c, const int *blocks, int nblocks, int nlev, int ncycles,
                                 const std::vector<std::vector<BlockRec *>> &lv, double *resi)
{
    int rc = pool_ready(c);
    if (rc) return rc;
    for (int q = 0; q < ff_ctx::NPOOL; q++)
        if (!c->mb_part[q] && hipMalloc(&c->mb_part[q], sizeof(double) * 7 * c->red_parts) != hipSuccess)
            return fail(c, FF_E_NOMEM, "RESL2 partials of the stream pool%s");
    if (c->mb_out_n < nblocks) {
        if (c->mb_out) (void)hipFree(c->mb_out);
        c->mb_out = nullptr;
        c->mb_out_n = 0;
        if (hipMalloc(&c->mb_out, sizeof(double) * 8 * nblocks) != hipSuccess)
            return fail(c, FF_E_NOMEM, "RESL2 sums of the blocks%s");
        c->mb_out_n = nblocks;
    }
    if (staged_mode(c))
        for (auto &l : lv)
            for (BlockRec *r : l)
                if ((rc = staged_alloc(c, r))) return rc;
    const hipStream_t origin = c->stream;
    const int np = std::min(nblocks, (int)ff_ctx::NPOOL);
    double dummy[7];
    for (int it = 0; it < ncycles && !rc; it++) {
        const bool sums = resi && it == ncycles - 1;
        HIPCHK(c, hipEventRecord(c->pool_ev[ff_ctx::NPOOL], origin));
        for (int q = 0; q < np; q++) HIPCHK(c, hipStreamWaitEvent(c->pool[q], c->pool_ev[ff_ctx::NPOOL], 0));
        c->in_pool = true;
        for (int n = 0; n < nblocks && !rc; n++) {
            const int q = n % np;
            c->stream = c->pool[q];
            const RedBuf rb{c->mb_part[q], c->red_parts, c->mb_out + 8 * n};
            rc = vcycle_once(c, lv[n], nlev, blocks[n], sums ? dummy : nullptr, &rb, false);
        }
        c->in_pool = false;
        c->stream = origin;
        for (int q = 0; q < np; q++) {
            if (hipEventRecord(c->pool_ev[q], c->pool[q]) != hipSuccess ||
                hipStreamWaitEvent(origin, c->pool_ev[q], 0) != hipSuccess)
                return fail(c, FF_E_HIP, "ff_vcycle_blocks: join of the stream pool failed%s");
        }
        if (rc) return rc;
        if ((rc = halo_impl(c, 1, 0))) return rc;
    }
    if (resi && ncycles > 0) {
        std::vector<double> h((size_t)8 * nblocks);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->mb_out, sizeof(double) * 8 * nblocks, hipMemcpyDeviceToHost, origin));
        HIPCHK(c, hipStreamSynchronize(origin));
        const int nq = c->prm.iturb == 6 ? 7 : 5;
        double sum[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int n = 0; n < nblocks; n++)
            for (int q = 0; q < 7; q++) sum[q] += q < nq ? h[(size_t)8 * n + q] : 0.;
        for (int q = 0; q < 7; q++) resi[q] = sum[q];
    } else if (resi) {
        for (int q = 0; q < 7; q++) resi[q] = 0.;
    }
    return FF_OK;
}

extern "C" int ff_vcycle_blocks(ff_ctx *c, const int *blocks, int nblocks, int nlev, int ncycles,
                                double *resi)
{
    if (!c || !blocks || nblocks < 1 || nlev < 1 || ncycles < 0) return FF_E_ARG;
    std::vector<std::vector<BlockRec *>> lv(nblocks);
    for (int n = 0; n < nblocks; n++) {
        int rc = vcycle_levels(c, blocks[n], nlev, lv[n]);
        if (rc) return rc;
    }
    if (c->block_streams && nblocks > 1) {
        int rc = vcycle_blocks_streams(c, blocks, nblocks, nlev, ncycles, lv, resi);
        if (rc) return rc;
        if (c->lusgs_mode != FF_LUSGS_PLANES)
            for (int n = 0; n < nblocks; n++)
                for (int m = 0; m < nlev; m++)
                    if ((rc = sweep_check(c, *lv[n][m]))) return rc;
        return FF_OK;
    }
    double sum[7] = {0, 0, 0, 0, 0, 0, 0}, rs[7];
    for (int it = 0; it < ncycles; it++) {
        const bool last = it == ncycles - 1;
        for (int n = 0; n < nblocks; n++) {
            int rc = vcycle_once(c, lv[n], nlev, blocks[n], last && resi ? rs : nullptr);
            if (rc) return rc;
            if (last && resi)
                for (int q = 0; q < 7; q++) sum[q] += rs[q];
        }
        int rc = halo_impl(c, 1, 0);
        if (rc) return rc;
    }
    if (resi)
        for (int q = 0; q < 7; q++) resi[q] = sum[q];
    if (c->lusgs_mode != FF_LUSGS_PLANES)
        for (int n = 0; n < nblocks; n++)
            for (int m = 0; m < nlev; m++) {
                int rc = sweep_check(c, *lv[n][m]);
                if (rc) return rc;
            }
    return FF_OK;
}

// The multi-block cycle as one HIP graph.  Every block's V-cycle is captured on a stream of
// a pool of eight (block n on stream n mod 8), forked from and joined back to the context's
// stream through events, so that the graph holds up to eight independent branches that the
// device runs side by side -- small blocks no longer leave most of the chip idle between
// their launches -- followed by the level-1 halo exchange.  Replays make no host launches.
extern "C" int ff_vcycle_blocks_graph(ff_ctx *c, const int *blocks, int nblocks, int nlev,
                                      int ncycles)
{
    if (!c || !blocks || nblocks < 1 || nlev < 1 || ncycles < 0) return FF_E_ARG;
    if (c->profile) return fail(c, FF_E_STATE, "ff_vcycle_blocks_graph: per-kernel timing is on%s");
    if (!c->stream) return fail(c, FF_E_STATE, "ff_vcycle_blocks_graph: the null stream cannot be captured%s");
    // A capture of 64 blocks of 4 levels (8 V-cycles of 4 levels on each of the eight
    // branches) ended the process with a segmentation fault while it was captured or first
    // replayed; 24 blocks of 3 levels (3 V-cycles per branch) replay correctly.  Until the
    // cause is known, captures beyond 3 blocks x 3 levels per branch are refused.
    if ((nblocks + ff_ctx::NBRANCH - 1) / ff_ctx::NBRANCH * nlev > 9)
        return fail(c, FF_E_UNSUP, "ff_vcycle_blocks_graph: more than 9 block levels per branch "
                                   "(3 blocks x 3 levels) are not captured%s");
    std::vector<std::vector<BlockRec *>> lv(nblocks);
    for (int n = 0; n < nblocks; n++) {
        int rc = vcycle_levels(c, blocks[n], nlev, lv[n]);
        if (rc) return rc;
    }
    std::vector<long long> key = {nlev, c->lusgs_mode, c->residual_mode, c->earsm_mode,
                                  c->precision, c->gen};
    key.insert(key.end(), blocks, blocks + nblocks);
    if (!c->mb_exec || key != c->mb_key) {
        if (c->mb_exec) (void)hipGraphExecDestroy(c->mb_exec);
        if (c->mb_graph) (void)hipGraphDestroy(c->mb_graph);
        c->mb_exec = nullptr;
        c->mb_graph = nullptr;
        if (int e = pool_ready(c)) return e;
        if (staged_mode(c))
            for (auto &l : lv)
                for (BlockRec *r : l) {
                    int rc = staged_alloc(c, r);
                    if (rc) return rc;
                }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (auto &l : lv)
            for (BlockRec *r : l)
                if (r->walls.pending) {
                    HIPCHK(c, hipEventSynchronize(r->walls.done));
                    r->walls.pending = false;
                }
        const hipStream_t origin = c->stream;
        HIPCHK(c, hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
        int rc = FF_OK;
        const int np = std::min(nblocks, (int)ff_ctx::NBRANCH);
        if (hipEventRecord(c->pool_ev[ff_ctx::NPOOL], origin) != hipSuccess) rc = FF_E_HIP;
        for (int q = 0; q < np && !rc; q++)
            if (hipStreamWaitEvent(c->pool[q], c->pool_ev[ff_ctx::NPOOL], 0) != hipSuccess) rc = FF_E_HIP;
        for (int n = 0; n < nblocks && !rc; n++) {
            c->stream = c->pool[n % np];
            rc = vcycle_once(c, lv[n], nlev, blocks[n], nullptr);
            // the block's last force pass joins its branch
            for (BlockRec *r : lv[n])
                if (r->walls.pending) {
                    if (!rc && hipStreamWaitEvent(c->stream, r->walls.done, 0) != hipSuccess) rc = FF_E_HIP;
                    r->walls.pending = false;
                }
        }
        c->stream = origin;
        for (int q = 0; q < np && !rc; q++)
            if (hipEventRecord(c->pool_ev[q], c->pool[q]) != hipSuccess ||
                hipStreamWaitEvent(origin, c->pool_ev[q], 0) != hipSuccess)
                rc = FF_E_HIP;
        if (!rc) rc = halo_impl(c, 1, 0);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(origin, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc == FF_E_HIP ? fail(c, FF_E_HIP, "ff_vcycle_blocks_graph: capture failed%s") : rc;
        }
        if (ec != hipSuccess || !g) return fail(c, FF_E_HIP, "%s", hipGetErrorString(ec));
        c->mb_graph = g;
        if (hipGraphInstantiate(&c->mb_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            c->mb_graph = nullptr;
            return fail(c, FF_E_HIP, "ff_vcycle_blocks_graph: instantiation failed%s");
        }
        c->mb_key = key;
    }
    for (auto &l : lv)
        for (BlockRec *r : l)
            if (r->walls.pending) {
                HIPCHK(c, hipStreamWaitEvent(c->stream, r->walls.done, 0));
                r->walls.pending = false;
            }
    for (int it = 0; it < ncycles; it++) HIPCHK(c, hipGraphLaunch(c->mb_exec, c->stream));
    if (c->lusgs_mode != FF_LUSGS_PLANES)
        for (auto &l : lv)
            for (BlockRec *r : l) {
                int rc = sweep_check(c, *r);
                if (rc) return rc;
            }
    return FF_OK;
}
