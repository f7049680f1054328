// ffconv.hip — SUBIN's convergence monitor on the device: the largest change of the cycle
// (DROMXX res3c.f:9293-9324 on the change SUBIN forms, ns3c.f:9836-9905).
// Included at the end of ffhip.hip (it uses the context and launch helpers defined there).
//
// DROMXX scans the level-1 slab in storage order and keeps a running maximum DROMAX with
// a tolerance: entry L replaces it only if |X(L)| > |DROMAX|*1.000001 and, after the ghost
// weighting (x .99999 per ghost direction), > |DROMAX|.  Which entry wins therefore depends
// on the order, and the result is reproduced exactly by a serial pass -- but only over the
// entries that can win.  With P_L the largest weighted magnitude before L, the running
// maximum before L is at least P_L / 1.000001 (the entry that attained P_L either became the
// maximum or was turned away by a maximum within that factor), so L can win only if its
// weighted magnitude a_L > P_L / 1.000001.  The device forms the prefix maxima in parallel
// (chunk maxima, their exclusive prefix, a block scan inside each chunk), keeps in order the
// entries with a_L * 1.0000011 >= P_L (a slightly wider test, so rounding can only add
// candidates) and a_L > the incoming maximum, and one wave runs DROMXX's statements over
// those candidates alone.

constexpr int DMX_T = 256;                 // threads per chunk workgroup
constexpr int DMX_PER = 8;                 // consecutive entries per thread
constexpr int DMX_CH = DMX_T * DMX_PER;    // entries per chunk

// the change of entry n that IDRXX selects (SUBIN ns3c.f:9819, 9840-9844)
__device__ __forceinline__ double dmx_change(const Blk &b, int idrxx, int n)
{
    switch (idrxx) {
    case 2: return FLD(b, FF_RO)[n] * (FLD(b, FF_U)[n] - FLD(b, FF_UOLD)[n]);
    case 3: return FLD(b, FF_RO)[n] * (FLD(b, FF_V)[n] - FLD(b, FF_VOLD)[n]);
    case 4: return FLD(b, FF_RO)[n] * (FLD(b, FF_W)[n] - FLD(b, FF_WOLD)[n]);
    case 5: return FLD(b, FF_TEMP)[n] - FLD(b, FF_TOLD)[n];
    case 6: return FLD(b, FF_PDIFF)[n] - FLD(b, FF_POLD)[n];
    default: return FLD(b, FF_REPS)[n] - FLD(b, FF_EPSOLD)[n];
    }
}

// the ghost weighting of DROMXX (res3c.f:9303-9306, with IJKPAI multi.f:3236): .99999 per
// direction in which entry n lies in a ghost layer, applied I, J, K in turn
__device__ __forceinline__ double dmx_weight(const Blk &b, int n, double x)
{
    const int k = n / b.il, r = n - k * b.il, j = r / b.is, i = r - j * b.is;
    if (i < FF_GHOST || i >= b.imax + FF_GHOST) x = x * .99999;
    if (j < FF_GHOST || j >= b.jmax + FF_GHOST) x = x * .99999;
    if (k < FF_GHOST || k >= b.kmax + FF_GHOST) x = x * .99999;
    return x;
}

// weighted magnitudes of chunk c, staged in LDS in entry order (coalesced loads)
__device__ __forceinline__ void dmx_stage(const Blk &b, int idrxx, int c, double *sa)
{
#pragma unroll
    for (int k = 0; k < DMX_PER; k++) {
        const int e = k * DMX_T + threadIdx.x, n = c * DMX_CH + e;
        sa[e] = n < b.ntot ? dmx_weight(b, n, fabs(dmx_change(b, idrxx, n))) : 0.;
    }
}

// phase 1: the largest weighted magnitude of every chunk
__global__ void __launch_bounds__(DMX_T) k_dmx_chunkmax(Blk b, int idrxx, double *cmax)
{
    __shared__ double sa[DMX_CH];
    __shared__ double sh[DMX_T];
    dmx_stage(b, idrxx, blockIdx.x, sa);
    double m = 0.;
#pragma unroll
    for (int k = 0; k < DMX_PER; k++) m = dmax(m, sa[k * DMX_T + threadIdx.x]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = DMX_T / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = dmax(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) cmax[blockIdx.x] = sh[0];
}

// inclusive max-scan of one value per thread of the workgroup (DMX_T threads)
__device__ __forceinline__ double dmx_block_scan_max(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = 1; h < DMX_T; h <<= 1) {
        const double o = threadIdx.x >= h ? sh[threadIdx.x - h] : 0.;
        __syncthreads();
        sh[threadIdx.x] = dmax(sh[threadIdx.x], o);
        __syncthreads();
    }
    return sh[threadIdx.x];
}

// inclusive sum-scan of one count per thread of the workgroup
__device__ __forceinline__ int dmx_block_scan_sum(int v, int *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = 1; h < DMX_T; h <<= 1) {
        const int o = threadIdx.x >= h ? sh[threadIdx.x - h] : 0;
        __syncthreads();
        sh[threadIdx.x] += o;
        __syncthreads();
    }
    return sh[threadIdx.x];
}

// phase 2: exclusive prefix maximum of the chunk maxima (one workgroup, each thread a run
// of consecutive chunks)
__global__ void __launch_bounds__(DMX_T) k_dmx_prefix(const double *cmax, int nch, double *pin)
{
    __shared__ double sh[DMX_T];
    const int per = (nch + DMX_T - 1) / DMX_T, c0 = threadIdx.x * per;
    double m = 0.;
    for (int c = c0; c < c0 + per && c < nch; c++) m = dmax(m, cmax[c]);
    dmx_block_scan_max(m, sh);
    double run = threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0.;
    for (int c = c0; c < c0 + per && c < nch; c++) {
        pin[c] = run;
        run = dmax(run, cmax[c]);
    }
}

// phase 3: the candidates of each chunk, in entry order, and their count
__global__ void __launch_bounds__(DMX_T) k_dmx_cand(Blk b, int idrxx, const double *pin,
                                                     const double *state, int *cand, int *cnt)
{
    __shared__ double sa[DMX_CH];
    __shared__ double shd[DMX_T];
    __shared__ int shi[DMX_T];
    const int c = blockIdx.x;
    dmx_stage(b, idrxx, c, sa);
    __syncthreads();
    const double m_in = fabs(state[0]);
    // this thread's entries DMX_PER*t .. DMX_PER*t + DMX_PER - 1 of the chunk
    double v[DMX_PER];
    double tmax = 0.;
#pragma unroll
    for (int k = 0; k < DMX_PER; k++) {
        v[k] = sa[threadIdx.x * DMX_PER + k];
        tmax = dmax(tmax, v[k]);
    }
    dmx_block_scan_max(tmax, shd);
    double p = threadIdx.x > 0 ? shd[threadIdx.x - 1] : 0.;
    p = dmax(p, pin[c]);
    unsigned flags = 0;
    int nc = 0;
#pragma unroll
    for (int k = 0; k < DMX_PER; k++) {
        const int n = c * DMX_CH + threadIdx.x * DMX_PER + k;
        if (n < b.ntot && v[k] > m_in && v[k] * 1.0000011 >= p) {
            flags |= 1u << k;
            nc++;
        }
        p = dmax(p, v[k]);
    }
    const int end = dmx_block_scan_sum(nc, shi);
    int pos = end - nc;
#pragma unroll
    for (int k = 0; k < DMX_PER; k++)
        if (flags & (1u << k)) cand[(size_t)c * DMX_CH + pos++] = c * DMX_CH + threadIdx.x * DMX_PER + k;
    if (threadIdx.x == DMX_T - 1) cnt[c] = end;
}

// phase 4: DROMXX's statements over the candidates in entry order (one wave; the lanes
// fetch 64 candidates' changes at a time, lane 0 runs the recurrence); state = DROMAX,
// IXERR (1-based), INERR in and out
__global__ void __launch_bounds__(64) k_dmx_final(Blk b, int idrxx, const int *cand,
                                                  const int *cnt, int nch, int nblock,
                                                  double *state)
{
    __shared__ double xs[64];
    __shared__ int ns[64];
    __shared__ int cns[64];
    const int lane = threadIdx.x;
    double dm = state[0];
    int ix = (int)state[1], inr = (int)state[2];
    for (int c0 = 0; c0 < nch; c0 += 64) {
        const int myc = c0 + lane;
        cns[lane] = myc < nch ? cnt[myc] : 0;
        __syncthreads();
        for (int q = 0; q < 64 && c0 + q < nch; q++) {
            const int nq = cns[q];
            for (int base = 0; base < nq; base += 64) {
                if (base + lane < nq) {
                    const int n = cand[(size_t)(c0 + q) * DMX_CH + base + lane];
                    ns[lane] = n;
                    xs[lane] = dmx_change(b, idrxx, n);
                }
                __syncthreads();
                if (lane == 0) {
                    const int m = min(64, nq - base);
                    for (int r = 0; r < m; r++) {
                        double apu = fabs(xs[r]);
                        if (apu > fabs(dm) * 1.000001) {
                            apu = dmx_weight(b, ns[r], apu);
                            if (apu > fabs(dm)) {
                                ix = ns[r] + 1;
                                inr = nblock;
                                dm = apu;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        if (nblock == inr) dm = dmx_weight(b, ix - 1, dmx_change(b, idrxx, ix - 1));
        state[0] = dm;
        state[1] = (double)ix;
        state[2] = (double)inr;
    }
}

extern "C" int ff_dromax(ff_ctx *c, int block, int idrxx, double *dromax, int *ixerr, int *inerr)
{
    if (!c || !dromax || !ixerr || !inerr) return FF_E_ARG;
    if (idrxx < 2 || idrxx > 8 || idrxx == 7)
        return fail(c, FF_E_UNSUP, "ff_dromax: IDRXX %s (2-6 and 8 supported)", std::to_string(idrxx).c_str());
    BlockRec *r = find(c, 1, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    const Blk &b = r->dev;
    // the entry IXERR of this block is read back when INERR names it and no entry wins
    if (*inerr == block && (*ixerr < 1 || *ixerr > b.ntot))
        return fail(c, FF_E_ARG, "ff_dromax: IXERR outside block %s", std::to_string(block).c_str());
    const int nch = (b.ntot + DMX_CH - 1) / DMX_CH;
    if (!r->dmx_mem) {
        const size_t bytes = sizeof(int) * ((size_t)nch * DMX_CH + nch) + sizeof(double) * (2 * (size_t)nch + 4);
        if (hipMalloc(&r->dmx_mem, bytes) != hipSuccess) return fail(c, FF_E_NOMEM, "DROMAX scratch%s");
    }
    double *cmax = (double *)r->dmx_mem, *pin = cmax + nch, *state = pin + nch;
    int *cnt = (int *)(state + 4), *cand = cnt + nch;
    c->lev = 1;
    c->red_host[0] = *dromax;
    c->red_host[1] = (double)*ixerr;
    c->red_host[2] = (double)*inerr;
    HIPCHK(c, hipMemcpyAsync(state, c->red_host, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = launch(c, "k_dmx_chunkmax", k_dmx_chunkmax, dim3(nch), dim3(DMX_T), b, idrxx, cmax);
    if (!rc) rc = launch(c, "k_dmx_prefix", k_dmx_prefix, dim3(1), dim3(DMX_T), (const double *)cmax, nch, pin);
    if (!rc) rc = launch(c, "k_dmx_cand", k_dmx_cand, dim3(nch), dim3(DMX_T), b, idrxx,
                         (const double *)pin, (const double *)state, cand, cnt);
    if (!rc) rc = launch(c, "k_dmx_final", k_dmx_final, dim3(1), dim3(64), b, idrxx,
                         (const int *)cand, (const int *)cnt, nch, block, state);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->red_host, state, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dromax = c->red_host[0];
    *ixerr = (int)c->red_host[1];
    *inerr = (int)c->red_host[2];
    return FF_OK;
}
