/*
 * ORACLE (test infrastructure only) -- type-generic body of the CPU
 * restatement of GridInfCuda's min-sum BP.  Included twice by bp_oracle.c,
 * once with OT=float/OS=f32 and once with OT=int/OS=i32, mirroring the two
 * explicit instantiations of the reference (src/bp.cc:190-194).
 *
 * Layouts follow the reference Matrix<T> (src/matrix.cc:16-19, :52-56):
 * the first index is fastest.
 *   data   [l][y][x]        = Matrix<T>(3, w, h, L)      (src/bp.cc:96)
 *   msgs   [l][d][y][x]     = Matrix<T>(4, w, h, 4, L)   (src/bp.cc:125)
 *   belief [l][y][x]        = Matrix<T>(3, w, h, L)      (src/bp.cc:126)
 *   table  V(a,b) at a + L*b = Matrix<T>(2, L, L)       (src/grid.cc:8, :59-61)
 *   labels [y][x]           = Matrix<int>(2, w, h)
 * msgs holds OUTBOXES: msgs(x,y,d,l) is the message node (x,y) sends to its
 * neighbour in direction d (src/bp.cc:155-161).
 */

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define OFN(name) OCAT(name, OS)

static inline size_t OFN(d_idx)(int w, int h, int x, int y, int l) {
    return (size_t)x + (size_t)w * ((size_t)y + (size_t)h * (size_t)l);
}

static inline size_t OFN(m_idx)(int w, int h, int x, int y, int d, int l) {
    return (size_t)x + (size_t)w * ((size_t)y + (size_t)h * ((size_t)d + 4u * (size_t)l));
}

/* Data-cost pyramid, one level (src/bp.cc:100-117).  coarse is (cw x ch x L)
 * with cw = ceil(w/2), ch = ceil(h/2); it starts at zero and every fine node,
 * visited in row-major order, adds into its parent, so a parent sums its
 * children in the order (2X,2Y), (2X+1,2Y), (2X,2Y+1), (2X+1,2Y+1). */
void OFN(oracle_pyramid)(const OT *fine, int w, int h, int L, OT *coarse) {
    int cw = (w + 1) / 2, ch = (h + 1) / 2;
    memset(coarse, 0, sizeof(OT) * (size_t)cw * ch * L);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int l = 0; l < L; l++)
                coarse[OFN(d_idx)(cw, ch, x / 2, y / 2, l)] += fine[OFN(d_idx)(w, h, x, y, l)];
}

/* Coarse-to-fine message initialisation in the OUTBOX frame
 * (src/bp.cc:129-139): every fine node, border included, copies the
 * outboxes of its parent (x/2, y/2). */
void OFN(oracle_upsample)(const OT *cmsg, int cw, int ch, OT *fmsg, int w, int h, int L) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int d = 0; d < 4; d++)
                for (int l = 0; l < L; l++)
                    fmsg[OFN(m_idx)(w, h, x, y, d, l)] = cmsg[OFN(m_idx)(cw, ch, x / 2, y / 2, d, l)];
}

/* bpCollect (src/bp.cc:7-18): b(l) = D(l) + in_S(l) + in_E(l) + in_N(l) + in_W(l),
 * added left to right, where in_d is the outbox of the neighbour p+dir(d)
 * that points back at p (direction OPP(d)). */
static void OFN(collect)(const OT *data, const OT *msgs, OT *belief, int w, int h, int L, int x, int y) {
    for (int l = 0; l < L; l++) {
        OT b = data[OFN(d_idx)(w, h, x, y, l)];
        for (int d = 0; d < 4; d++)
            b += msgs[OFN(m_idx)(w, h, x + DIR_X[d], y + DIR_Y[d], OPP_DIR(d), l)];
        belief[OFN(d_idx)(w, h, x, y, l)] = b;
    }
}

/* bpSendMsg (src/bp.cc:21-63) for node (x,y) towards direction d, result in out[L]. */
static void OFN(send_msg)(const OT *msgs, const OT *belief, int w, int h, int L, int model,
                          const OT *table, int norm, int x, int y, int d, OT *out) {
    int nx = x + DIR_X[d], ny = y + DIR_Y[d];
    if (model == ORACLE_SM_TRUNC_LINEAR || model == ORACLE_SM_POTTS) {
        /* O(L) distance transform (src/bp.cc:26-54). */
        OT minh = 0;
        for (int l = 0; l < L; l++) {
            out[l] = belief[OFN(d_idx)(w, h, x, y, l)] - msgs[OFN(m_idx)(w, h, nx, ny, OPP_DIR(d), l)];
            minh = (l == 0) ? out[0] : OMIN(minh, out[l]);
        }
        /* s = V(0,1), tau = V(0,L-1) (src/bp.cc:35, :43).  With L == 1 the
         * scans are empty and s is never used. */
        OT s = (L > 1) ? table[0 + L * 1] : (OT)0;
        OT tau = table[0 + L * (L - 1)];
        for (int l = 1; l < L; l++) out[l] = OMIN(out[l], out[l - 1] + s);
        for (int l = L - 2; l >= 0; l--) out[l] = OMIN(out[l], out[l + 1] + s);
        minh += tau;
        OT acc = 0;
        for (int l = 0; l < L; l++) {
            out[l] = OMIN(out[l], minh);
            acc += out[l];
        }
        OT nu;
        if (norm == ORACLE_NORM_MEAN) {
            /* Mean of the clamped message in T arithmetic (src/bp.cc:50-51):
             * correctly rounded float division, or int division truncating
             * toward zero. */
            nu = acc / (OT)L;
        } else {
            /* Min normalisation (project extension: the message minimum is 0). */
            nu = out[0];
            for (int l = 1; l < L; l++) nu = OMIN(nu, out[l]);
        }
        for (int l = 0; l < L; l++) out[l] -= nu;
    } else {
        /* General O(L^2) path for EXPLICIT / TRUNC_QUAD (src/bp.cc:55-62):
         * no reverse-message subtraction and no normalisation. */
        for (int i = 0; i < L; i++) {
            OT v = table[i + L * 0] + belief[OFN(d_idx)(w, h, x, y, 0)];
            for (int j = 0; j < L; j++) {
                OT c = table[i + L * j] + belief[OFN(d_idx)(w, h, x, y, j)];
                v = OMIN(v, c);
            }
            out[i] = v;
        }
    }
}

/* One red-black half-sweep number `it` on one level (src/bp.cc:148-165):
 * every interior node with (x + y + it) odd collects, then sends its four
 * messages, which overwrite its outboxes.  Nodes of one colour are
 * independent, so rows may run in parallel with identical results. */
void OFN(oracle_half_sweep)(const OT *data, OT *msgs, OT *belief, int w, int h, int L, int model,
                            const OT *table, int norm, int it, int threads) {
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
    {
        OT *out = (OT *)malloc(sizeof(OT) * (size_t)L);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int y = 1; y < h - 1; y++) {
            for (int x = ((y + it) & 1) + 1; x < w - 1; x += 2) {
                OFN(collect)(data, msgs, belief, w, h, L, x, y);
                for (int d = 0; d < 4; d++) {
                    OFN(send_msg)(msgs, belief, w, h, L, model, table, norm, x, y, d, out);
                    for (int l = 0; l < L; l++) msgs[OFN(m_idx)(w, h, x, y, d, l)] = out[l];
                }
            }
        }
        free(out);
    }
    (void)threads;
}

/* Stale-belief labelling (src/bp.cc:65-84, :168-173): for interior nodes,
 * the lowest label minimising the belief stored at the node's last update
 * (zero if it was never updated).  Border labels are left untouched. */
void OFN(oracle_labels)(const OT *belief, int w, int h, int L, int *labels) {
    for (int y = 1; y < h - 1; y++)
        for (int x = 1; x < w - 1; x++) {
            OT best = belief[OFN(d_idx)(w, h, x, y, 0)];
            int arg = 0;
            for (int l = 1; l < L; l++) {
                OT c = belief[OFN(d_idx)(w, h, x, y, l)];
                if (c < best) { best = c; arg = l; }
            }
            labels[(size_t)y * w + x] = arg;
        }
}

/* decodeHbp (src/bp.cc:86-180); decodeBp is levels == 1 (src/bp.cc:182-187).
 * msgs0_out / belief0_out (optional) receive the level-0 state at the end.
 * Returns 0, or a negative code for arguments the reference leaves undefined. */
int OFN(oracle_hbp)(const OT *data, int W, int H, int L, int model, const OT *table, int levels,
                    int iters, int norm, int *labels, OT *msgs0_out, OT *belief0_out, int threads) {
    if (W < 1 || H < 1 || L < 1 || levels < 1 || iters < 0) return -1;
    OT **dl = (OT **)calloc((size_t)levels, sizeof(OT *));
    int *lw = (int *)calloc((size_t)levels, sizeof(int));
    int *lh = (int *)calloc((size_t)levels, sizeof(int));
    lw[0] = W; lh[0] = H;
    dl[0] = (OT *)malloc(sizeof(OT) * (size_t)W * H * L);
    memcpy(dl[0], data, sizeof(OT) * (size_t)W * H * L);
    for (int t = 1; t < levels; t++) {
        lw[t] = (lw[t - 1] + 1) / 2;
        lh[t] = (lh[t - 1] + 1) / 2;
        dl[t] = (OT *)malloc(sizeof(OT) * (size_t)lw[t] * lh[t] * L);
        OFN(oracle_pyramid)(dl[t - 1], lw[t - 1], lh[t - 1], L, dl[t]);
    }
    OT *msgs = NULL, *belief = NULL;
    for (int t = levels - 1; t >= 0; t--) {
        int w = lw[t], h = lh[t];
        OT *m = (OT *)calloc((size_t)w * h * 4 * L, sizeof(OT));
        OT *b = (OT *)calloc((size_t)w * h * L, sizeof(OT));
        if (msgs) {
            OFN(oracle_upsample)(msgs, lw[t + 1], lh[t + 1], m, w, h, L);
            free(msgs);
            free(belief);
        }
        msgs = m; belief = b;
        for (int it = 0; it < iters; it++)
            OFN(oracle_half_sweep)(dl[t], msgs, belief, w, h, L, model, table, norm, it, threads);
    }
    OFN(oracle_labels)(belief, W, H, L, labels);
    if (msgs0_out) memcpy(msgs0_out, msgs, sizeof(OT) * (size_t)W * H * 4 * L);
    if (belief0_out) memcpy(belief0_out, belief, sizeof(OT) * (size_t)W * H * L);
    free(msgs); free(belief);
    for (int t = 0; t < levels; t++) free(dl[t]);
    free(dl); free(lw); free(lh);
    return 0;
}

/* Grid::getLabelingCost (src/grid.cc:102-125), accumulated in T in the
 * reference order: interior nodes, row-major, each adding D(f), V(f, f_N),
 * V(f, f_E). */
OT OFN(oracle_energy)(const OT *data, const OT *table, const int *labels, int W, int H, int L) {
    OT e = 0;
    for (int y = 1; y < H - 1; y++)
        for (int x = 1; x < W - 1; x++) {
            int f = labels[(size_t)y * W + x];
            int fn = labels[(size_t)(y - 1) * W + x];
            int fe = labels[(size_t)y * W + x + 1];
            e += data[OFN(d_idx)(W, H, x, y, f)];
            e += table[f + L * fn];
            e += table[f + L * fe];
        }
    return e;
}

/* The same term set accumulated in double (the FP64 energy). */
double OFN(oracle_energy_f64)(const OT *data, const OT *table, const int *labels, int W, int H, int L) {
    double e = 0.0;
    for (int y = 1; y < H - 1; y++)
        for (int x = 1; x < W - 1; x++) {
            int f = labels[(size_t)y * W + x];
            int fn = labels[(size_t)(y - 1) * W + x];
            int fe = labels[(size_t)y * W + x + 1];
            e += (double)data[OFN(d_idx)(W, H, x, y, f)];
            e += (double)table[f + L * fn];
            e += (double)table[f + L * fe];
        }
    return e;
}

/* Grid::getLabelingCost(f, x, y, label) (src/grid.cc:127-136): D(label) and
 * then V(label, f(p + dir d)) for d = S, E, N, W, accumulated in T. */
static OT OFN(local_cost)(const OT *data, const OT *table, const int *f, int W, int H, int L, int x, int y,
                          int label) {
    OT c = data[OFN(d_idx)(W, H, x, y, label)];
    for (int d = 0; d < 4; d++) c += table[label + L * f[(size_t)(y + DIR_Y[d]) * W + x + DIR_X[d]]];
    return c;
}

/* Lowest-index label of minimal local cost (strict <, src/icm.cc:18-25). */
static int OFN(icm_best)(const OT *data, const OT *table, const int *f, int W, int H, int L, int x, int y) {
    OT best = OFN(local_cost)(data, table, f, W, H, L, x, y, 0);
    int arg = 0;
    for (int i = 1; i < L; i++) {
        OT c = OFN(local_cost)(data, table, f, W, H, L, x, y, i);
        if (c < best) { best = c; arg = i; }
    }
    return arg;
}

/* decodeIcmSync (src/icm.cc:5-37): every interior node moves to its best
 * label given the previous iteration's labelling; borders keep `initial`. */
void OFN(oracle_icm_sync)(const OT *data, const OT *table, int W, int H, int L, const int *initial, int iters,
                          int *result) {
    size_t n = (size_t)W * H;
    int *old = (int *)malloc(n * sizeof(int));
    memcpy(result, initial, n * sizeof(int));
    memcpy(old, result, n * sizeof(int));
    for (int t = 0; t < iters; t++) {
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) result[(size_t)y * W + x] = OFN(icm_best)(data, table, old, W, H, L, x, y);
        memcpy(old, result, n * sizeof(int));
    }
    free(old);
}

/* decodeIcmAsync (src/icm.cc:39-62): the same update in place, row-major. */
void OFN(oracle_icm_async)(const OT *data, const OT *table, int W, int H, int L, const int *initial, int iters,
                           int *result) {
    memcpy(result, initial, (size_t)W * H * sizeof(int));
    for (int t = 0; t < iters; t++)
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++)
                result[(size_t)y * W + x] = OFN(icm_best)(data, table, result, W, H, L, x, y);
}

/* decodeMaxPrior (src/include/estimators.h:10-25): every node (borders
 * included) takes the lowest-index label of minimal data cost. */
void OFN(oracle_max_prior)(const OT *data, int W, int H, int L, int *result) {
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int m = 0;
            for (int i = 1; i < L; i++)
                if (data[OFN(d_idx)(W, H, x, y, i)] < data[OFN(d_idx)(W, H, x, y, m)]) m = i;
            result[(size_t)y * W + x] = m;
        }
}

#undef OFN
#undef OCAT
#undef OCAT2
