// kernels.hpp -- HIP kernels of the step path (gfx950, wave64).
//
// Layout: beads live in "slots"; every replica occupies a 64-aligned slot
// range, so no wavefront spans two replicas (the RNG jump-ahead and the
// per-replica reductions are wave-uniform).  Neighbour rows are sliced
// ELL: entry k of bead s sits at slice_start[s/64] + 64*k + s%64, so a
// wavefront reads a row column with one coalesced load.
//
// Step sequence (parity p = step & 1; positions ping-pong pos[p] ->
// pos[1-p] because neighbours read old positions while owners integrate):
//   k_qa_pairs    sift free pairs + advance contacts      (qa/*.cpp)
//   k_qa_serial   accept candidates, apply removals       (process_candidates.cpp)
//   k_force       all force terms per bead (+ integrate)  (forces/*.cpp, lang_pc.cpp)
//   k_integrate   when PID forces spill to neighbours
//   k_finalize    RNG/time advance, verify decision       (nl/verify.cpp)
//   k_cell_* / k_nl_* / k_fp_write  Verlet rebuild        (nl/legacy_update.cpp + update_X)
#pragma once
#include "common.hpp"

namespace xmd_k {

// ------------------------------------------------------------ RNG jumps ---
// y = M x over GF(2), M given by 64 rows (lane r holds row r): every lane
// ends with the same y.  Needs a full wavefront.
__device__ inline uint64_t wave_matvec_rows(const uint64_t *rows, uint64_t x) {
    uint64_t row = rows[lane_id()];
    int bit = __popcll(row & x) & 1;
    return (uint64_t)__ballot(bit);
}
// State of the wave's lane l at draw offset 4*(64*w + l) from S (uniform).
// The wave offset is one mat-vec with J^(256 w) (table jwave, w < jwave_n;
// beyond it, by the binary powers in jrow); the lane offset J^(4l) applied
// to the uniform state is an XOR of 16 table entries, one per nibble of
// the state: jnib[(16 k + v) 64 + l] = J^(4l) applied to v << 4k.  All 16
// loads are independent.
__device__ inline uint64_t wave_lane_state(const Dev &D, uint64_t S, int w) {
    if (w < D.jwave_n) S = wave_matvec_rows(D.jwave + 64 * (size_t)w, S);
    else
        for (int j = 0; w != 0; ++j, w >>= 1)
            if (w & 1) S = wave_matvec_rows(D.jrow + 64 * j, S);
    const int l = lane_id();
    uint64_t y = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) y ^= D.jnib[64 * ((k << 4) | (int)((S >> (4 * k)) & 15u)) + l];
    return y;
}

// -------------------------------------------------------------- QA pairs ---
// Chunks [0, nfp_chunks): sift 256 free pairs each (qa/sift_candidates.cpp);
// chunks after that: process 256 contacts each (qa/process_contacts.cpp).
// Candidates and removals are compacted per chunk in index order, so the
// serial pass sees them in set order without sorting.  Grid-stride over the
// live chunks only (those below each replica's live counts).
__global__ __launch_bounds__(BLK) void k_qa_pairs(Dev D, Derived G, int p, int mutate) {
    // s_pre (2 nrep + 1 ints, dynamic): prefix over the segments "free pairs
    // of replica r" (r < nrep, only when sifting) and "contacts of replica
    // r" of their live chunk counts; block wi of the work sits in segment g
    // with s_pre[g] <= wi < s_pre[g+1].
    extern __shared__ int s_pre[];
    __shared__ int wsum[BLK / WAVE];
    const int tid = threadIdx.x;
    const int nseg = 2 * D.nrep;
    int nwork = 0;
    for (int g0 = 0; g0 < nseg; g0 += BLK) {
        const int g = g0 + tid;
        int live = 0;
        if (g < D.nrep) live = mutate ? (D.fp_count[g] + BLK - 1) / BLK : 0;
        else if (g < nseg) live = (D.ct_count[g - D.nrep] + BLK - 1) / BLK;
        int x = live;
        for (int o = 1; o < WAVE; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane_id() >= o) x += y;
        }
        if (lane_id() == WAVE - 1) wsum[tid / WAVE] = x;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < BLK / WAVE; ++k) {
            if (k < tid / WAVE) woff += wsum[k];
            tot += wsum[k];
        }
        if (g < nseg) s_pre[g + 1] = nwork + woff + x;
        __syncthreads();
        nwork += tot;
    }
    if (tid == 0) s_pre[0] = 0;
    __syncthreads();
    const double *P = D.pos[p];
    const int ct_tot = D.ct_cap[D.nrep];
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        int lo = 0, hi = nseg - 1;  // largest g with s_pre[g] <= wi (a non-empty segment)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pre[mid] <= wi) lo = mid;
            else hi = mid - 1;
        }
        // the segment names the replica and the chunk's place in it, so
        // its base, live count and box load together
        const bool sift = lo < D.nrep;
        if (sift && !mutate) continue;
        const int rep = sift ? lo : lo - D.nrep;
        if (mutate && D.err[rep]) continue;  // halted replica
        const int first = (wi - s_pre[lo]) * BLK;
        const int count = sift ? D.fp_count[rep] : D.ct_count[rep];
        const int chunk = (sift ? D.fp_base[rep] : D.ct_base[rep]) / BLK + (wi - s_pre[lo]);
        if (first >= count) continue;
        int idx = chunk * BLK + tid;
        int local = first + tid;
        Box box = D.box[rep];
        bool flag = false;
        int type = 0;
        uint32_t d1 = 0, d2 = 0;
        if (sift) {
            // entries below the live count hold valid slots even when vacant,
            // so the pair, its vacancy flag and its positions load without
            // waiting on the flag
            if (local < count) {
                const int i1 = D.fp_i1[idx], i2 = D.fp_i2[idx];
                const uint8_t vac = D.fp_vac[idx];
                v3<double> r1 = ldp(P, i1), r2 = ldp(P, i2);
                v3<double> r12 = r_uv(box, r1, r2);
                // qa_sift rejects a pair of type k when 1 > r12_rn req[k]; with
                // req[k] <= req_max that holds for every k once it holds for
                // req_max (rounding is monotone), so far pairs skip n and h
                if (!vac && !(1.0 > norm_inv(r12) * G.qa.req_max)) {
                    uint8_t f1 = D.bflags[i1], f2 = D.bflags[i2];
                    bool ok1 = (f1 & B_PREV) && (f1 & B_NEXT), ok2 = (f2 & B_PREV) && (f2 & B_NEXT);
                    v3<double> n1, h1, n2, h2;
                    qa_nh(ok1, ok1 ? ldp(P, i1 - 1) : r1, r1, ok1 ? ldp(P, i1 + 1) : r1, n1, h1);
                    qa_nh(ok2, ok2 ? ldp(P, i2 - 1) : r2, r2, ok2 ? ldp(P, i2 + 1) : r2, n2, h2);
                    type = qa_sift(G.qa, D.req_min_dist, D.ptype, r12, n1, h1, n2, h2, D.atype[i1], D.atype[i2],
                                   D.sync[i1], D.sync[i2], d1, d2);
                    flag = type != 0;
                }
            }
        } else {
            if (local < count && !D.ct_vac[idx]) {
                int i1 = D.ct_i1[idx], i2 = D.ct_i2[idx];
                v3<double> r12 = r_uv(box, ldp(P, i1), ldp(P, i2));
                double depth, rmin, rmax;
                int ctype = D.ct_type[idx], status = D.ct_status[idx], next;
                qa_lj(G.qa, D.ss_depth, D.ss_rmin, D.ss_rmax, ctype, depth, rmin, rmax);
                double t = D.tt[p][rep];
                FV<double> f = qa_contact(G.qa, depth, rmin, rmax, r12, t, D.ct_ref[idx], status, next);
                D.ct_F[idx] = f.F.x;
                D.ct_F[ct_tot + idx] = f.F.y;
                D.ct_F[2 * ct_tot + idx] = f.F.z;
                D.ct_V[idx] = f.V;
                if (mutate) {
                    if (next == 1 && status == 0) {
                        D.ct_status[idx] = 1;
                        D.ct_ref[idx] = t;
                    }
                    flag = next == 2;
                }
            }
            if (!mutate) continue;
        }
        // ordered compaction of the chunk
        unsigned long long m = __ballot(flag);
        int w = tid / WAVE;
        if (lane_id() == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = 0;
        for (int k = 0; k < w; ++k) off += wsum[k];
        int rank = off + __popcll(m & ((1ull << lane_id()) - 1ull));
        int base = chunk * BLK;
        if (flag) {
            if (sift) {
                D.cand_fp[base + rank] = idx;
                D.cand_type[base + rank] = type;
                D.cand_d1[base + rank] = d1;
                D.cand_d2[base + rank] = d2;
            } else {
                D.rm_idx[base + rank] = idx;
            }
        }
        if (tid == 0) {
            int tot = 0;
            for (int k = 0; k < BLK / WAVE; ++k) tot += wsum[k];
            if (sift) D.cand_cnt[chunk] = tot;
            else D.rm_cnt[chunk] = tot;
        }
        __syncthreads();
    }
}

// Contact slot lists of a bead: slot_add_q appends ci at position q (the
// bead's current count), slot_remove takes ci out keeping the order.
__device__ inline void slot_add_q(const Dev &D, int s, int q, int ci) {
    if (q >= MAXQ) {
        atomicOr(D.err + D.slot_rep[s], E_SLOT);
        return;
    }
    D.qslot[q * D.nslot + s] = ci;
    D.qn[s] = q + 1;
}
__device__ inline void slot_remove(const Dev &D, int s, int ci) {
    int q = D.qn[s], k = 0;
    while (k < q && D.qslot[k * D.nslot + s] != ci) ++k;
    if (k == q) return;
    for (; k + 1 < q; ++k) D.qslot[k * D.nslot + s] = D.qslot[(k + 1) * D.nslot + s];
    D.qn[s] = q - 1;
}
// One contact removal (process_contacts.cpp): the pair's budgets return to
// its beads, the pair rejoins the free pairs at position o of the
// replica's set (or flags the overflow), and leaves its beads' slot lists.
__device__ inline void remove_contact(const Dev &D, int rep, int ci, int i1, int i2, int fpb, int o, int fcap) {
    D.ct_vac[ci] = 1;
    D.sync[i1] = sync_add(D.sync[i1], D.ct_d1[ci]);
    D.sync[i2] = sync_add(D.sync[i2], D.ct_d2[ci]);
    if (o >= fcap) {
        atomicOr(D.err + rep, E_FP_CAP);
    } else {
        const int f = fpb + o;
        D.fp_i1[f] = i1;
        D.fp_i2[f] = i2;
        D.fp_vac[f] = 0;
    }
    slot_remove(D, i1, ci);
    slot_remove(D, i2, ci);
}

// List, in order, the chunks [from, c_hi) whose counter is non-zero:
// s_list[q] is the chunk, s_off[q] the sum of the counters of the chunks
// listed before it (s_off[n] the total).  LIST_PER consecutive chunks per
// thread, their counters loaded together.  Returns the number listed n
// and sets *next to the first chunk not examined.
constexpr int LIST_PER = 8, LIST_SPAN = LIST_PER * BLK;
__device__ inline int list_nonzero_chunks(const int *cnt, int c_hi, int from, int *s_list, int *s_off, int *next) {
    __shared__ int s_w[2][BLK / WAVE];
    const int tid = threadIdx.x, c0 = from + LIST_PER * tid;
    int v[LIST_PER];
#pragma unroll
    for (int i = 0; i < LIST_PER; ++i) v[i] = c0 + i < c_hi ? cnt[c0 + i] : 0;
    int mine = 0, ments = 0;
#pragma unroll
    for (int i = 0; i < LIST_PER; ++i) {
        mine += v[i] > 0;
        ments += v[i];
    }
    // exclusive scans of the per-thread chunk and entry counts: wave
    // shuffles, then waves
    int x = mine, y = ments;
    for (int o = 1; o < WAVE; o <<= 1) {
        const int xo = __shfl_up(x, o), yo = __shfl_up(y, o);
        if (lane_id() >= o) {
            x += xo;
            y += yo;
        }
    }
    const int w = tid / WAVE;
    if (lane_id() == WAVE - 1) {
        s_w[0][w] = x;
        s_w[1][w] = y;
    }
    __syncthreads();
    int off = x - mine, eoff = y - ments, tot = 0, etot = 0;
    for (int k = 0; k < BLK / WAVE; ++k) {
        if (k < w) {
            off += s_w[0][k];
            eoff += s_w[1][k];
        }
        tot += s_w[0][k];
        etot += s_w[1][k];
    }
#pragma unroll
    for (int i = 0; i < LIST_PER; ++i)
        if (v[i] > 0) {
            s_list[off] = c0 + i;
            s_off[off++] = eoff;
            eoff += v[i];
        }
    if (tid == 0) s_off[tot] = etot;
    __syncthreads();
    *next = from + LIST_SPAN;
    return tot;
}
// The listed chunk holding entry k (0 <= k < s_off[n]): the largest q with
// s_off[q] <= k (listed chunks are non-empty, so the offsets increase).
__device__ inline int listed_chunk(const int *s_off, int n, int k) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_off[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One batch of candidates or removals staged in LDS.
struct QaBatch {
    int e[BLK], i1[BLK], i2[BLK], ty[BLK], q1[BLK], q2[BLK], dup[BLK];
    uint32_t d1[BLK], d2[BLK], y1[BLK], y2[BLK];
};

// One workgroup per replica: process candidates in free-pair order, then
// the removals of process_contacts in contact order, then compact the
// contact set when its extent passes half the capacity.  The workgroup
// lists the non-empty chunks and stages their entries; the order-dependent
// steps run on one thread.
__global__ __launch_bounds__(BLK) void k_qa_serial(Dev D, int p) {
    __shared__ int s_scan[BLK];
    __shared__ int s_list[LIST_SPAN], s_off[LIST_SPAN + 1];
    __shared__ QaBatch s_b;
    __shared__ int s_state[4];  // ct count, fp count, cnt0
    int rep = blockIdx.x, tid = threadIdx.x;
    if (D.err[rep]) return;  // halted replica: its QA state stays as it is
    int fpb = D.fp_base[rep], ctb = D.ct_base[rep];
    int fc_lo = fpb / BLK, fc_hi = fc_lo + (D.fp_count[rep] + BLK - 1) / BLK;
    double t = D.tt[p][rep];
    if (tid == 0) {
        s_state[0] = D.ct_count[rep];
        s_state[1] = D.fp_count[rep];
        s_state[2] = s_state[0];
    }
    __syncthreads();
    const int cap = D.ct_cap[rep], tot = D.ct_cap[D.nrep], fcap = D.fp_cap[rep];
    // candidates, in free-pair order.  The entries of the listed chunks
    // are staged in batches of BLK (pair, budgets, slot counts, loaded by
    // the whole block), then one thread accepts them in order from LDS; a
    // bead that an earlier candidate of the batch names is read back from
    // memory, where that thread wrote it.  Once the contact set is full
    // every later acceptance fails alike (the error is set, nothing else
    // changes).
    for (int from = fc_lo; from < fc_hi;) {
        int nxt;
        const int nl = list_nonzero_chunks(D.cand_cnt, fc_hi, from, s_list, s_off, &nxt);
        const int ne = s_off[nl];
        for (int b0 = 0; b0 < ne; b0 += BLK) {
            const int k = b0 + tid, nb = min(BLK, ne - b0);
            int i1 = -1, i2 = -1;
            if (k < ne) {
                const int q = listed_chunk(s_off, nl, k);
                const int e = s_list[q] * BLK + (k - s_off[q]);
                const int fp = D.cand_fp[e];
                i1 = D.fp_i1[fp];
                i2 = D.fp_i2[fp];
                s_b.e[tid] = fp;
                s_b.i1[tid] = i1;
                s_b.i2[tid] = i2;
                s_b.ty[tid] = D.cand_type[e];
                s_b.d1[tid] = D.cand_d1[e];
                s_b.d2[tid] = D.cand_d2[e];
                s_b.y1[tid] = D.sync[i1];
                s_b.y2[tid] = D.sync[i2];
                s_b.q1[tid] = D.qn[i1];
                s_b.q2[tid] = D.qn[i2];
            }
            __syncthreads();
            if (k < ne) {
                int dup = 0;
                for (int m = 0; m < tid; ++m) {
                    const int a = s_b.i1[m], b = s_b.i2[m];
                    dup |= (a == i1 || b == i1) ? 1 : 0;
                    dup |= (a == i2 || b == i2) ? 2 : 0;
                }
                s_b.dup[tid] = dup;
            }
            __syncthreads();
            if (tid == 0) {
                int cnt = s_state[0];
                for (int m = 0; m < nb; ++m) {
                    const int a1 = s_b.i1[m], a2 = s_b.i2[m], dup = s_b.dup[m];
                    const uint32_t d1 = s_b.d1[m], d2 = s_b.d2[m];
                    const uint32_t ns1 = sync_sub(dup & 1 ? D.sync[a1] : s_b.y1[m], d1);
                    const uint32_t ns2 = sync_sub(dup & 2 ? D.sync[a2] : s_b.y2[m], d2);
                    if (!(sync_valid(ns1) && sync_valid(ns2))) continue;
                    if (cnt >= cap) {
                        atomicOr(D.err + rep, E_CT_CAP);
                        continue;
                    }
                    const int ci = ctb + cnt++;
                    D.ct_i1[ci] = a1;
                    D.ct_i2[ci] = a2;
                    D.ct_type[ci] = s_b.ty[m];
                    D.ct_status[ci] = 0;
                    D.ct_ref[ci] = t;
                    D.ct_d1[ci] = d1;
                    D.ct_d2[ci] = d2;
                    D.ct_vac[ci] = 0;
                    D.ct_F[ci] = 0;
                    D.ct_F[tot + ci] = 0;
                    D.ct_F[2 * tot + ci] = 0;
                    D.ct_V[ci] = 0;
                    D.sync[a1] = ns1;
                    D.sync[a2] = ns2;
                    slot_add_q(D, a1, dup & 1 ? D.qn[a1] : s_b.q1[m], ci);
                    slot_add_q(D, a2, dup & 2 ? D.qn[a2] : s_b.q2[m], ci);
                    D.fp_vac[s_b.e[m]] = 1;
                }
                s_state[0] = cnt;
            }
            __syncthreads();
        }
        from = nxt;
    }
    // removals, in contact order (contacts formed above have none), in
    // batches of BLK: a removal whose beads no other removal of the batch
    // names runs on its own thread; the others run in order on one thread.
    // Removal o of the pass (counted from its start) puts its pair back at
    // free-pair position fp_count + o, or flags the overflow.
    int cc_lo = ctb / BLK, cc_hi = cc_lo + (s_state[2] + BLK - 1) / BLK;
    for (int from = cc_lo; from < cc_hi;) {
        int nxt;
        const int nl = list_nonzero_chunks(D.rm_cnt, cc_hi, from, s_list, s_off, &nxt);
        const int ne = s_off[nl];
        for (int b0 = 0; b0 < ne; b0 += BLK) {
            const int k = b0 + tid, nb = min(BLK, ne - b0);
            const int fc = s_state[1];
            int ci = -1, i1 = -1, i2 = -1;
            if (k < ne) {
                const int q = listed_chunk(s_off, nl, k);
                ci = D.rm_idx[s_list[q] * BLK + (k - s_off[q])];
                i1 = D.ct_i1[ci];
                i2 = D.ct_i2[ci];
                s_b.e[tid] = ci;
                s_b.i1[tid] = i1;
                s_b.i2[tid] = i2;
            }
            __syncthreads();
            if (k < ne) {
                bool shared = false;
                for (int m = 0; m < nb; ++m) {
                    const int a = s_b.i1[m], b = s_b.i2[m];
                    shared |= m != tid && (a == i1 || b == i1 || a == i2 || b == i2);
                }
                s_b.dup[tid] = shared;
                if (!shared) remove_contact(D, rep, ci, i1, i2, fpb, fc + tid, fcap);
            }
            __syncthreads();
            if (tid == 0) {
                for (int m = 0; m < nb; ++m)
                    if (s_b.dup[m]) remove_contact(D, rep, s_b.e[m], s_b.i1[m], s_b.i2[m], fpb, fc + m, fcap);
                s_state[1] = min(fcap, fc + nb);
            }
            __syncthreads();
        }
        for (int q = tid; q < nl; q += BLK) D.rm_cnt[s_list[q]] = 0;
        __syncthreads();
        from = nxt;
    }
    if (tid == 0) {
        D.fp_count[rep] = s_state[1];
        D.ct_count[rep] = s_state[0];
        s_scan[0] = s_state[0] > cap / 2;
    }
    __syncthreads();
    if (!s_scan[0]) return;
    __syncthreads();
    // compaction: in-place, chunk by chunk in index order (new <= old)
    int cnt = s_state[0];
    int kept = 0;
    for (int c0 = 0; c0 < cnt; c0 += BLK) {
        int o = c0 + tid;
        bool keep = o < cnt && !D.ct_vac[ctb + o];
        int i1 = 0, i2 = 0, ty = 0, st = 0;
        double rf = 0, fx = 0, fy = 0, fz = 0, v = 0;
        uint32_t a = 0, b = 0;
        if (keep) {
            int ci = ctb + o;
            i1 = D.ct_i1[ci]; i2 = D.ct_i2[ci]; ty = D.ct_type[ci]; st = D.ct_status[ci];
            rf = D.ct_ref[ci]; a = D.ct_d1[ci]; b = D.ct_d2[ci];
            fx = D.ct_F[ci]; fy = D.ct_F[tot + ci]; fz = D.ct_F[2 * tot + ci]; v = D.ct_V[ci];
        }
        s_scan[tid] = keep;
        __syncthreads();
        for (int sh = 1; sh < BLK; sh <<= 1) {
            int x = tid >= sh ? s_scan[tid - sh] : 0;
            __syncthreads();
            s_scan[tid] += x;
            __syncthreads();
        }
        int nw = kept + s_scan[tid] - keep;
        int total = s_scan[BLK - 1];
        __syncthreads();
        if (keep) {
            int ci = ctb + nw;
            D.ct_i1[ci] = i1; D.ct_i2[ci] = i2; D.ct_type[ci] = ty; D.ct_status[ci] = st;
            D.ct_ref[ci] = rf; D.ct_d1[ci] = a; D.ct_d2[ci] = b; D.ct_vac[ci] = 0;
            D.ct_F[ci] = fx; D.ct_F[tot + ci] = fy; D.ct_F[2 * tot + ci] = fz; D.ct_V[ci] = v;
            // remember old -> new in rm_idx's space (old index slot)
            D.rm_idx[ctb + o] = ci;
        }
        kept += total;
        __syncthreads();
    }
    // remap bead slots
    int s0 = D.rep_slot_off[rep], s1 = s0 + D.rep_n[rep];
    for (int s = s0 + tid; s < s1; s += BLK) {
        int q = D.qn[s];
        for (int k = 0; k < q; ++k) {
            int old = D.qslot[k * D.nslot + s];
            D.qslot[k * D.nslot + s] = D.rm_idx[old];
        }
    }
    __syncthreads();
    for (int o = kept + tid; o < cnt; o += BLK) D.ct_vac[ctb + o] = 1;
    if (tid == 0) D.ct_count[rep] = kept;
}

// Langevin predictor-corrector of one bead, xmd/src/dynamics/lang_pc.cpp:11-37,
// plus the displacement test of nl/verify.cpp:10-15 (max over the replica).
__device__ inline void integrate_bead_impl(const Dev &D, const Derived &G, int p, int s, int rep, bool active, double Fx,
                                           double Fy, double Fz) {
    const int n = D.nslot;
    double *y = D.y;
    // the bead's state first: these loads overlap the noise's dependent
    // chain (generator state -> jump tables -> log, sin, cos -> the
    // correction table for the few draws near a rounding boundary)
    v3<double> y0{0, 0, 0}, y1{0, 0, 0}, y2{0, 0, 0}, y3{0, 0, 0}, y4{0, 0, 0}, y5{0, 0, 0}, r_orig{0, 0, 0};
    if (active) {
        y0 = ldp(D.pos[p], s);
        y1 = ld3(y, n, s), y2 = ld3(y + 3 * n, n, s), y3 = ld3(y + 6 * n, n, s);
        y4 = ld3(y + 9 * n, n, s), y5 = ld3(y + 12 * n, n, s);
        r_orig = ldp(D.orig_r, s);
    }
    const double nsd = D.noise_sd[rep];
    const Box box = D.box[rep];
    int local = s - D.rep_slot_off[rep];
    uint64_t S = wave_lane_state(D, D.rng[p][rep], local >> 6);
    float nz[3];
    bead_noise(S, NoiseFix{D.bm_fix, D.bm_thr}, nz);
    // a zero uniform in a log() slot gives infinite noise (rand_gen.h:45-46)
    if (active && (!((rng_out(S) >> 32)) || !((rng_out(rng_step(rng_step(S))) >> 32)))) atomicOr(D.err + rep, E_ZERO_DRAW);
    double disp = 0.0;
    if (active) {
        if (G.lang_on) {
            v3<double> v = y1 * G.dt_inv;
            v3<double> a;
            a.x = Fx * 1.0 - G.gamma * v.x + nsd * (double)nz[0];
            a.y = Fy * 1.0 - G.gamma * v.y + nsd * (double)nz[1];
            a.z = Fz * 1.0 - G.gamma * v.z + nsd * (double)nz[2];
            v3<double> err = y2 - a * G.half_dt2;
            y0 = y0 - (3.0 / 16.0) * err;
            y1 = y1 - (251.0 / 360.0) * err;
            y2 = y2 - 1.0 * err;
            y3 = y3 - (11.0 / 18.0) * err;
            y4 = y4 - (1.0 / 6.0) * err;
            y5 = y5 - (1.0 / 60.0) * err;
            y0 = y0 + ((((y1 + y2) + y3) + y4) + y5);
            y1 = y1 + (((2.0 * y2 + 3.0 * y3) + 4.0 * y4) + 5.0 * y5);
            y2 = y2 + ((3.0 * y3 + 6.0 * y4) + 10.0 * y5);
            y3 = y3 + (4.0 * y4 + 10.0 * y5);
            y4 = y4 + 5.0 * y5;
        }
        stp(D.pos[1 - p], s, y0);
        st3(y, n, s, y1);
        st3(y + 3 * n, n, s, y2);
        st3(y + 6 * n, n, s, y3);
        st3(y + 9 * n, n, s, y4);
        st3(y + 12 * n, n, s, y5);
        v3<double> dr = r_uv(box, y0, r_orig);
        disp = norm(dr);
        if (!(disp == disp) || !(fabs(y0.x) < 1e300)) disp = __builtin_nan("");
    }
    // wave max (non-negative doubles order like their bit patterns; NaN
    // wins), stored for the wavefront's slice: a plain store, where an
    // atomic per wavefront on one word per replica would serialise at the
    // memory side (the XCDs' L2s are not coherent)
    unsigned long long bits = active ? (unsigned long long)xmd_bm::dbits(disp) : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long other = __shfl_xor(bits, o);
        bits = other > bits ? other : bits;
    }
    if (lane_id() == 0) D.slice_maxd[p][s >> 6] = bits;
}

// ------------------------------------------------------------- forces ---
// Integration when PID spilled forces onto chain neighbours.
__global__ __launch_bounds__(BLK) void k_integrate(Dev D, Derived G, int p) {
    const int n = D.nslot;
    int s = blockIdx.x * BLK + threadIdx.x;
    int rep = D.slot_rep[s];
    uint8_t bf = D.bflags[s];
    bool active = bf & B_ACTIVE;
    // branch-free loads (in-bounds slots stand in for missing neighbours) so
    // that nothing waits for them before the state and noise loads issue;
    // the selects keep the sum order F + prev + next
    const bool hp = bf & B_PREV, hn = bf & B_NEXT;
    const int sp = hp ? s - 1 : s, sn = hn ? s + 1 : s;
    const double *ps = D.pid_side;
    const double fx = D.Fbuf[s], fy = D.Fbuf[n + s], fz = D.Fbuf[2 * n + s];
    const double px = ps[3 * n + sp], py = ps[4 * n + sp], pz = ps[5 * n + sp];
    const double qx = ps[sn], qy = ps[n + sn], qz = ps[2 * n + sn];
    double Fx = hp ? fx + px : fx, Fy = hp ? fy + py : fy, Fz = hp ? fz + pz : fz;
    Fx = hn ? Fx + qx : Fx;
    Fy = hn ? Fy + qy : Fy;
    Fz = hn ? Fz + qz : Fz;
    if (!active) Fx = Fy = Fz = 0.0;
    // a halted replica (error word set) keeps its state: only the
    // positions cross to the other parity buffer (uniform per wavefront:
    // replicas are 64-aligned)
    if (D.err[rep]) {
        if (active) stp(D.pos[1 - p], s, ldp(D.pos[p], s));
        return;
    }
    integrate_bead_impl(D, G, p, s, rep, active, Fx, Fy, Fz);
}

// Total force of the non-fused path (eval_forces): F += PID spill.
__global__ __launch_bounds__(BLK) void k_gather_force(Dev D) {
    const int n = D.nslot;
    int s = blockIdx.x * BLK + threadIdx.x;
    uint8_t bf = D.bflags[s];
    if (!(bf & B_ACTIVE)) return;
    const double *ps = D.pid_side;
    double Fx = D.Fbuf[s], Fy = D.Fbuf[n + s], Fz = D.Fbuf[2 * n + s];
    if (bf & B_PREV) {
        Fx += ps[3 * n + s - 1];
        Fy += ps[4 * n + s - 1];
        Fz += ps[5 * n + s - 1];
    }
    if (bf & B_NEXT) {
        Fx += ps[s + 1];
        Fy += ps[n + s + 1];
        Fz += ps[2 * n + s + 1];
    }
    D.Fbuf[s] = Fx;
    D.Fbuf[n + s] = Fy;
    D.Fbuf[2 * n + s] = Fz;
}

// One wavefront per replica: advance rand_gen by 4N draws and the time by dt
// (lang_pc.cpp:39-40), then decide the Verlet rebuild (verify.cpp:17-21).
// mode 0: decision from force_rb only (no advance).  Returns the decision
// to every lane.
__device__ inline int finalize_wave(const Dev &D, const Derived &G, int p, int mode, int rep) {
    const int l = lane_id();
    if (rep == 0 && l == 0) D.any_rb[1 - p] = 0;
    if (mode == 1 && D.err[rep]) {
        // halted replica: generator and time carried over unchanged, no
        // rebuild (a pending forced rebuild waits for the reset)
        if (l == 0) {
            D.rng[1 - p][rep] = D.rng[p][rep];
            D.tt[1 - p][rep] = D.tt[p][rep];
            D.rb_now[rep] = 0;
        }
        return 0;
    }
    if (mode == 1) {
        uint64_t Sn = wave_matvec_rows(D.jadv + 64 * rep, D.rng[p][rep]);
        if (l == 0) {
            D.rng[1 - p][rep] = Sn;
            D.tt[1 - p][rep] = D.tt[p][rep] + G.dt;
            D.n_done[rep] += 1;
        }
    }
    // the replica's largest displacement: the maximum over its slices,
    // each written by the integrator of this step (a running replica's
    // every wavefront stores its slice)
    unsigned long long mb = 0ull;
    if (mode == 1) {
        const int sl0 = D.rep_slot_off[rep] >> 6, sl1 = (D.rep_slot_off[rep] + D.rep_n[rep] + WAVE - 1) >> 6;
        for (int t = sl0 + l; t < sl1; t += WAVE) {
            const unsigned long long v = D.slice_maxd[p][t];
            mb = v > mb ? v : mb;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long other = __shfl_xor(mb, o);
            mb = other > mb ? other : mb;
        }
    }
    int rb = 0;
    if (l == 0) {
        rb = D.force_rb[rep];
        if (mode == 1) {
            double m = xmd_bm::bitsd(mb);
            if (!(m == m)) atomicOr(D.err + rep, E_NONFINITE);
            double total = 2.0 * m + 0.0;
            if (total > G.pad) rb = 1;
        }
        D.rb_now[rep] = rb;
        D.force_rb[rep] = 0;
        if (rb) atomicOr(D.any_rb + p, 1);
    }
    return __shfl(rb, 0);
}
__global__ __launch_bounds__(WAVE) void k_finalize(Dev D, Derived G, int p, int mode) {
    finalize_wave(D, G, p, mode, blockIdx.x);
}

// ------------------------------------------------------------- rebuild ---
__device__ inline int cell_coord(double x, double L, int nc) {
    double w = x - L * floor(x / L);
    int c = (int)(w / L * (double)nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// One workgroup per replica: finalize_wave, then -- when the replica
// rebuilds -- its whole cell list: counts in LDS, exclusive scan, scatter
// (cell_start / cell_end, cell_beads, bead_cell), and the free-pair slice
// counters the row fill adds to set to zero.  The engine uses it when every
// replica has at most CELLS_LDS cells and FUSED_MAX_BEADS beads, so that an
// idle step's finalize and cell phase is one small launch; larger replicas
// take k_finalize, k_cell_count, k_seg_scan and k_cell_fill.
constexpr int CELLS_LDS = 15360;
constexpr int FUSED_MAX_BEADS = 32768;
__global__ __launch_bounds__(1024) void k_cells(Dev D, Derived G, const double *P, int p, int mode) {
    __shared__ int s_c[CELLS_LDS];
    __shared__ int s_wsum[1024 / WAVE];
    __shared__ int s_rb;
    const int rep = blockIdx.x, tid = threadIdx.x, wid = tid / WAVE, l = lane_id();
    if (wid == 0) {
        const int rb = finalize_wave(D, G, p, mode, rep);
        if (l == 0) s_rb = rb;
    }
    __syncthreads();
    if (!s_rb) return;
    const int a = D.rep_slot_off[rep], b = a + D.rep_n[rep];
    for (int sl = (a >> 6) + tid; sl < ((b + WAVE - 1) >> 6); sl += 1024) D.fp_slice_cnt[sl] = 0;
    const int c0 = D.cell_off[rep], nc = D.cell_off[rep + 1] - c0;
    if (nc == 0) return;  // open box: the row fill scans the replica
    const int *nd = D.cell_dim + 3 * rep;
    const Box box = D.box[rep];
    for (int c = tid; c < nc; c += 1024) s_c[c] = 0;
    __syncthreads();
    for (int s = a + tid; s < b; s += 1024) {
        if (!(D.bflags[s] & B_ACTIVE)) continue;
        const v3<double> r = ldp(P, s);
        const int cx = cell_coord(r.x, box.cx, nd[0]), cy = cell_coord(r.y, box.cy, nd[1]),
                  cz = cell_coord(r.z, box.cz, nd[2]);
        const int c = (cz * nd[1] + cy) * nd[0] + cx;
        D.bead_cell[s] = c0 + c;
        atomicAdd(&s_c[c], 1);
    }
    __syncthreads();
    // exclusive scan in chunks of 1024 cells; cell_start is the first
    // entry of the cell in cell_beads (absolute: the replica's beads start
    // at its first slot)
    int carry = 0;
    for (int c00 = 0; c00 < nc; c00 += 1024) {
        const int c = c00 + tid;
        const int v = c < nc ? s_c[c] : 0;
        int x = v;
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (l >= o) x += y;
        }
        if (l == WAVE - 1) s_wsum[wid] = x;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < 1024 / WAVE; ++k) {
            const int t = s_wsum[k];
            if (k < wid) woff += t;
            tot += t;
        }
        if (c < nc) {
            const int ex = carry + woff + x - v;
            s_c[c] = ex;
            D.cell_start[c0 + c] = a + ex;
        }
        carry += tot;
        __syncthreads();
    }
    for (int s = a + tid; s < b; s += 1024) {
        if (!(D.bflags[s] & B_ACTIVE)) continue;
        const int k = atomicAdd(&s_c[D.bead_cell[s] - c0], 1);
        D.cell_beads[a + k] = s;
    }
    __syncthreads();
    for (int c = tid; c < nc; c += 1024) D.cell_end[c0 + c] = a + s_c[c];
}

__global__ __launch_bounds__(BLK) void k_cell_count(Dev D, const double *P, int p) {
    if (!D.any_rb[p]) return;
    int s = blockIdx.x * BLK + threadIdx.x;
    int rep = D.slot_rep[s];
    if (!D.rb_now[rep]) return;
    if ((s & (WAVE - 1)) == 0) D.fp_slice_cnt[s >> 6] = 0;  // the row fill adds to it
    if (!(D.bflags[s] & B_ACTIVE)) return;
    int c0 = D.cell_off[rep];
    if (D.cell_off[rep + 1] == c0) return;
    const int *nd = D.cell_dim + 3 * rep;
    Box b = D.box[rep];
    const v3<double> r = ldp(P, s);
    int cx = cell_coord(r.x, b.cx, nd[0]), cy = cell_coord(r.y, b.cy, nd[1]), cz = cell_coord(r.z, b.cz, nd[2]);
    int c = c0 + (cz * nd[1] + cy) * nd[0] + cx;
    D.bead_cell[s] = c;
    atomicAdd(D.cell_cnt + c, 1);
}

// One workgroup per replica: exclusive scan of the replica's cell counts
// (cell_cnt -> cell_start / cell_end / cell_cursor, absolute slot offsets).
__global__ __launch_bounds__(1024) void k_seg_scan(Dev D, int p) {
    __shared__ int s_buf[1024];
    __shared__ int s_carry;
    if (!D.any_rb[p]) return;
    int rep = blockIdx.x, tid = threadIdx.x;
    if (!D.rb_now[rep]) return;
    const int lo = D.cell_off[rep], hi = D.cell_off[rep + 1], base = D.rep_slot_off[rep];
    if (tid == 0) s_carry = base;
    __syncthreads();
    for (int c0 = lo; c0 < hi; c0 += 1024) {
        int c = c0 + tid;
        int v = c < hi ? D.cell_cnt[c] : 0;
        s_buf[tid] = v;
        __syncthreads();
        for (int sh = 1; sh < 1024; sh <<= 1) {
            int x = tid >= sh ? s_buf[tid - sh] : 0;
            __syncthreads();
            s_buf[tid] += x;
            __syncthreads();
        }
        int ex = s_carry + s_buf[tid] - v;
        if (c < hi) {
            D.cell_start[c] = ex;
            D.cell_end[c] = ex + v;
            D.cell_cursor[c] = ex;
        }
        __syncthreads();
        if (tid == 1023) s_carry = ex + v;
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLK) void k_cell_fill(Dev D, int p) {
    if (!D.any_rb[p]) return;
    int s = blockIdx.x * BLK + threadIdx.x;
    int rep = D.slot_rep[s];
    if (!(D.bflags[s] & B_ACTIVE) || !D.rb_now[rep]) return;
    if (D.cell_off[rep + 1] == D.cell_off[rep]) return;
    int c = D.bead_cell[s];
    int k = atomicAdd(D.cell_cursor + c, 1);
    D.cell_beads[k] = s;
}

// Row flags of the pair (s, j) at rebuild: legacy_update's criterion
// (legacy_update.cpp:24-29) and the per-term filters of update_X
// (pauli.cpp:13-20, es_pair.cpp:9-21, bundle.cpp:9-26, free_pair.cpp:9-19).
// The pair is a list candidate: not a chain neighbour within 2, closer than
// the list radius.
__device__ inline bool pair_near(const Dev &D, const Derived &G, const Box &box, int s, int j, v3<double> rs,
                                 const double *P) {
    if (D.chain[j] == D.chain[s] && (s > j ? s - j : j - s) < 3) return false;
    return norm(wrap(box, ldp(P, j) - rs)) < G.req_r;
}
// Flags of a pair that pair_near accepts.
__device__ inline uint8_t pair_flags_near(const Dev &D, const Derived &G, const Box &box, int s, int j, v3<double> rs,
                                          const double *P) {
    v3<double> d = wrap(box, ldp(P, j) - rs);
    double dist = norm(d);
    // disulfide pair: in native_ssbonds only, no pair term (D24)
    for (int k = D.ss_start[s]; k < D.ss_start[s + 1]; ++k)
        if (D.ss_j[k] == j) return F_NATIVE | F_NATSS;
    // native?  rows are short and sorted
    bool native = false;
    for (int k = D.nat_start[s]; k < D.nat_start[s + 1]; ++k)
        if (D.nat_j[k] == j) {
            native = true;
            break;
        }
    double ninv = 1.0 / dist;
    uint8_t f = native ? F_NATIVE : 0;
    if (G.pauli_on && !native && ninv > G.pauli_min_inv) f |= F_PAULI;
    if (G.es_on && ninv > G.es_min_inv && D.charge[D.atype[s]] * D.charge[D.atype[j]] != 0.0) f |= F_ES;
    if (G.pid_on && !native && ninv > G.pid_min_inv) {
        uint8_t a = D.bflags[s], b = D.bflags[j];
        if ((a & B_PREV) && (a & B_NEXT) && (b & B_PREV) && (b & B_NEXT)) f |= F_PID;
    }
    if (G.qa_on && !native && s < j && ninv > G.qa_min_inv) {
        bool in_contact = false;
        int q = D.qn[s];
        int tot = D.ct_cap[D.nrep];
        (void)tot;
        for (int k = 0; k < q; ++k) {
            int ci = D.qslot[k * D.nslot + s];
            if (D.ct_i1[ci] == s ? D.ct_i2[ci] == j : D.ct_i1[ci] == j) in_contact = true;
        }
        if (!in_contact) f |= F_QAFREE;
    }
    return f;
}
__device__ inline uint8_t pair_flags(const Dev &D, const Derived &G, const Box &box, int s, int j, v3<double> rs,
                                     const double *P) {
    return pair_near(D, G, box, s, j, rs, P) ? pair_flags_near(D, G, box, s, j, rs, P) : 0;
}

// Wavefront per slot: the lanes share the slot's candidate partners.  With
// cells, the 27 neighbour cells (dx fastest, then dy, dz) form one flat
// candidate sequence: lane c < 27 holds cell c's range, a wave scan gives
// the offsets, and each lane finds the cell of its candidate by bisection
// over them, so every iteration uses all 64 lanes.
template <typename Fn> __device__ inline void for_candidates_wave(const Dev &D, int s, int rep, Fn fn) {
    const int l = lane_id();
    int c0 = D.cell_off[rep];
    if (D.cell_off[rep + 1] == c0) {
        int a = D.rep_slot_off[rep], b = a + D.rep_n[rep];
        for (int j0 = a; j0 < b; j0 += WAVE) {
            int j = j0 + l;
            fn(j < b && j != s ? j : -1);
        }
        return;
    }
    const int *nd = D.cell_dim + 3 * rep;
    int c = D.bead_cell[s] - c0;
    int cx = c % nd[0], cy = (c / nd[0]) % nd[1], cz = c / (nd[0] * nd[1]);
    int start = 0, cnt = 0;
    if (l < 27) {
        int x = (cx + l % 3 - 1 + nd[0]) % nd[0], y = (cy + (l / 3) % 3 - 1 + nd[1]) % nd[1],
            z = (cz + l / 9 - 1 + nd[2]) % nd[2];
        int cc = c0 + (z * nd[1] + y) * nd[0] + x;
        start = D.cell_start[cc];
        cnt = D.cell_end[cc] - start;
    }
    int incl = cnt;
    for (int o = 1; o < 32; o <<= 1) {
        int v = __shfl_up(incl, o);
        if (l >= o) incl += v;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, 26);
    for (int k0 = 0; k0 < total; k0 += WAVE) {
        const int k = k0 + l;
        // largest cell q in [0, 27) with excl[q] <= k (excl[0] = 0)
        int lo = 0, hi = 26;
        for (int it = 0; it < 5; ++it) {
            int mid = (lo + hi + 1) >> 1;
            if (__shfl(excl, mid) <= k) lo = mid;
            else hi = mid - 1;
        }
        const int st = __shfl(start, lo), ex = __shfl(excl, lo);
        const int j = k < total ? D.cell_beads[st + k - ex] : -1;
        fn(j != s ? j : -1);
    }
}

constexpr int NL_WAVES = BLK / WAVE;  // slots per block in the rebuild kernels
constexpr int MAXSORT = 1024;  // longest row a slot may hold (the row width is at most this)

// Row length per slot (entries carrying any term flag); run once at engine
// creation to size the rows.  The rebuild
// kernels loop over slots with a grid of one resident round of blocks and
// skip replicas that do not rebuild this step (the test is uniform over a block: its NL_WAVES slots
// lie in one 64-aligned replica range), so a step where one replica
// rebuilds launches no idle wavefronts for the others.
__global__ __launch_bounds__(BLK) void k_nl_count(Dev D, Derived G, const double *P, int p) {
    if (!D.any_rb[p]) return;
    for (int s0 = blockIdx.x * NL_WAVES; s0 < D.nslot; s0 += gridDim.x * NL_WAVES) {
        const int s = s0 + threadIdx.x / WAVE;
        const int rep = D.slot_rep[s];
        if (!D.rb_now[rep]) continue;
        int cnt = 0;
        if (D.bflags[s] & B_ACTIVE) {
            const Box box = D.box[rep];
            const v3<double> rs = ldp(P, s);
            for_candidates_wave(D, s, rep, [&](int j) {
                if (j >= 0 && (pair_flags(D, G, box, s, j, rs, P) & (F_PAULI | F_ES | F_PID | F_QAFREE))) ++cnt;
            });
        }
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if (lane_id() == 0) D.nl_len[s] = cnt;
    }
}

// Gather each row in LDS (at most the row width; longer rows flag a
// capacity error), rank-sort it by partner (partners are distinct), write
// it; count the half row and the QA free pairs; snapshot orig_r and the Go
// activity.  Returns the slot's QA free pairs to every lane.
__device__ inline int nl_fill_slot(const Dev &D, const Derived &G, const double *P, int s, int rep, int *lj,
                                   uint8_t *lf) {
    const int l = lane_id();
    const bool active = D.bflags[s] & B_ACTIVE;
    const Box box = D.box[rep];
    const v3<double> rs = ldp(P, s);
    const int base = D.slice_start[s >> 6] + (s & 63);
    const int W = D.nl_w;
    int cnt = 0, nfp = 0;
    if (active) {
        // the candidates within the list radius first (most of the 27
        // cells' beads are not), compacted in candidate order into lj; then
        // their flags, the kept ones compacted in place (a chunk's entries
        // are read before any of its writes, which land below the next chunk)
        int nnear = 0;
        for_candidates_wave(D, s, rep, [&](int j) {
            const bool near = j >= 0 && pair_near(D, G, box, s, j, rs, P);
            const unsigned long long m = __ballot(near);
            const int pos = nnear + __popcll(m & ((1ull << l) - 1ull));
            if (near && pos < MAXSORT) lj[pos] = j;
            nnear += __popcll(m);
        });
        if (nnear > MAXSORT) {
            if (l == 0) {
                atomicOr(D.err + rep, E_NL_CAP);
                atomicMax(D.need_w + rep, nnear);
            }
            nnear = MAXSORT;
        }
        for (int k0 = 0; k0 < nnear; k0 += WAVE) {
            const int j = k0 + l < nnear ? lj[k0 + l] : -1;
            const uint8_t f = j >= 0 ? pair_flags_near(D, G, box, s, j, rs, P) : 0;
            const bool keep = f & (F_PAULI | F_ES | F_PID | F_QAFREE);
            const unsigned long long m = __ballot(keep);
            const int pos = cnt + __popcll(m & ((1ull << l) - 1ull));
            if (keep && pos < W) {
                lj[pos] = j;
                lf[pos] = f;
                nfp += (f & F_QAFREE) != 0;
            }
            cnt += __popcll(m);
        }
        if (cnt > W) {
            if (l == 0) {
                atomicOr(D.err + rep, E_NL_CAP);
                atomicMax(D.need_w + rep, cnt);
            }
            cnt = W;
        }
    }
    __syncthreads();  // reached by every wave of the block: the skip test is block-uniform
    if (active) {
        // the rank of each entry: every lane reads the same four entries per
        // LDS load (a broadcast)
        const int c4 = cnt & ~3;
        const int4 *lj4 = reinterpret_cast<const int4 *>(lj);
        for (int k = l; k < cnt; k += WAVE) {
            const int j = lj[k];
            int r = 0;
            for (int q = 0; q < c4; q += 4) {
                const int4 v = lj4[q >> 2];
                r += (v.x < j) + (v.y < j) + (v.z < j) + (v.w < j);
            }
            for (int q = c4; q < cnt; ++q) r += lj[q] < j;
            D.nl_e[base + 64 * r] = nle_pack(j, lf[k]);
        }
    }
    if (!active) {
        if (l == 0) {
            D.nl_len[s] = 0;
            D.fp_n[s] = 0;
        }
        return 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nfp += __shfl_xor(nfp, o);
    }
    if (l == 0) {
        D.nl_len[s] = cnt;
        D.fp_n[s] = nfp;
        stp(D.orig_r, s, rs);
    }
    // Go activity (go.cpp:11-22)
    for (int k = D.nat_start[s] + l; k < D.nat_start[s + 1]; k += WAVE) {
        int j = D.nat_j[k];
        v3<double> d = wrap(box, ldp(P, j) - rs);
        double cutoff = 2.0 * D.nat_d[k];
        D.nat_act[k] = norm(d) < cutoff + G.pad;
    }
    // current disulfide bonds (native_ss.cpp:8-19, D23)
    for (int k = D.ss_start[s] + l; k < D.ss_start[s + 1]; k += WAVE) {
        v3<double> d = wrap(box, ldp(P, D.ss_j[k]) - rs);
        D.ss_act[k] = norm(d) < G.ss_cut + G.pad;
    }
    return nfp;
}

__global__ __launch_bounds__(BLK) void k_nl_fill(Dev D, Derived G, const double *P, int p) {
    __shared__ __align__(16) int s_j[NL_WAVES][MAXSORT];
    __shared__ uint8_t s_f[NL_WAVES][MAXSORT];
    __shared__ int s_nfp[NL_WAVES];
    if (!D.any_rb[p]) return;
    const int w = threadIdx.x / WAVE;
    for (int s0 = blockIdx.x * NL_WAVES; s0 < D.nslot; s0 += gridDim.x * NL_WAVES) {
        const int s = s0 + w;
        const int rep = D.slot_rep[s];
        if (!D.rb_now[rep]) continue;  // uniform over the block
        const int nfp = nl_fill_slot(D, G, P, s, rep, s_j[w], s_f[w]);
        // the block's slots share a 64-slot slice: one add to its free-pair
        // count per block (the next iteration's writes of s_nfp follow the
        // barrier inside nl_fill_slot, which thread 0 reaches after its read)
        if (lane_id() == 0) s_nfp[w] = nfp;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int k = 0; k < NL_WAVES; ++k) t += s_nfp[k];
            if (t) atomicAdd(D.fp_slice_cnt + (s0 >> 6), t);
        }
    }
}

// Free pairs in canonical order (slot, then row order); zero the cell
// counters; count the rebuild.  Wavefront per slot: the slot's pairs follow
// those of the replica's lower slices (their counts, added by the row fill,
// summed here) and of the lower slots of its own 64-slot slice.  The wave of
// the first slot of the replica's last slice sets the replica's count and
// flags a capacity overflow (with the count it needed).
__global__ __launch_bounds__(BLK) void k_fp_write(Dev D, int p) {
    if (!D.any_rb[p]) return;
    const int l = lane_id();
    for (int s0 = blockIdx.x * NL_WAVES; s0 < D.nslot; s0 += gridDim.x * NL_WAVES) {
        const int s = s0 + threadIdx.x / WAVE;
        const int rep = D.slot_rep[s];
        if (!D.rb_now[rep]) continue;
        const int sl = s >> 6, sl0 = D.rep_slot_off[rep] >> 6;
        int pre = 0;
        for (int t = sl0 + l; t < sl; t += WAVE) pre += D.fp_slice_cnt[t];
        int before = l < (s & 63) ? D.fp_n[(sl << 6) + l] : 0;
        for (int o = 32; o > 0; o >>= 1) {
            pre += __shfl_xor(pre, o);
            before += __shfl_xor(before, o);
        }
        int off = D.fp_base[rep] + pre + before;
        const int end = D.fp_base[rep] + D.fp_cap[rep];
        const int base = D.slice_start[sl] + (s & 63);
        const int len = (D.bflags[s] & B_ACTIVE) && D.fp_n[s] ? D.nl_len[s] : 0;
        for (int k0 = 0; k0 < len; k0 += WAVE) {
            const int k = k0 + l, e = base + 64 * k;
            const uint32_t ent = k < len ? D.nl_e[e] : 0u;
            const bool keep = nle_flag(ent) & F_QAFREE;
            const unsigned long long m = __ballot(keep);
            const int o = off + __popcll(m & ((1ull << l) - 1ull));
            if (keep && o < end) {
                D.fp_i1[o] = s;
                D.fp_i2[o] = nle_j(ent);
                D.fp_vac[o] = 0;
            }
            off += __popcll(m);
        }
        if (l == 0) {
            const int local = s - D.rep_slot_off[rep];
            const int c0 = D.cell_off[rep], nc = D.cell_off[rep + 1] - c0;
            for (int c = local; c < nc; c += D.rep_n[rep]) D.cell_cnt[c0 + c] = 0;
            if (local == 0) D.n_rebuild[rep] += 1;
            if ((s & 63) == 0 && sl == ((D.rep_slot_off[rep] + D.rep_n[rep] - 1) >> 6)) {
                const int used = pre + D.fp_slice_cnt[sl];
                if (used > D.fp_cap[rep]) {
                    // the count the host grows the replica's region to (the
                    // overflow of the QA pass leaves it 0: that one halts)
                    atomicOr(D.err + rep, E_FP_CAP);
                    D.need_fp[rep] = used;
                }
                D.fp_count[rep] = used > D.fp_cap[rep] ? D.fp_cap[rep] : used;
            }
        }
    }
}

// ---------------------------------------------------------- observables ---
// Sum over a workgroup in FP64, fixed order (wave shuffle tree, then waves
// in order); every thread gets the total.
__device__ inline double block_sum(double v, double *lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < BLK / WAVE; ++k) t += lds[k];
    return t;
}

// Eigenvalues of a symmetric 3x3 matrix (a00, a11, a22, a01, a02, a12),
// ascending, by the trigonometric closed form of the characteristic cubic.
__device__ inline void sym3_eig(const double a[6], double e[3]) {
    const double p1 = a[3] * a[3] + a[4] * a[4] + a[5] * a[5];
    const double q = (a[0] + a[1] + a[2]) / 3.0;
    const double b0 = a[0] - q, b1 = a[1] - q, b2 = a[2] - q;
    const double p2 = b0 * b0 + b1 * b1 + b2 * b2 + 2.0 * p1;
    if (p2 <= 0.0) {
        e[0] = e[1] = e[2] = q;
        return;
    }
    const double pp = sqrt(p2 / 6.0), ip = 1.0 / pp;
    const double c0 = b0 * ip, c1 = b1 * ip, c2 = b2 * ip, c3 = a[3] * ip, c4 = a[4] * ip, c5 = a[5] * ip;
    double r = 0.5 * (c0 * (c1 * c2 - c5 * c5) - c3 * (c3 * c2 - c5 * c4) + c4 * (c3 * c5 - c1 * c4));
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    const double phi = acos(r) / 3.0;
    const double hi = q + 2.0 * pp * cos(phi), lo = q + 2.0 * pp * cos(phi + 2.0 * M_PI / 3.0);
    e[0] = lo;
    e[1] = 3.0 * q - hi - lo;
    e[2] = hi;
}

// Per-replica observables of the stats report (xmd/src/io/report_stats.cpp:
// 8-41): out[8 rep + k] = t, V (last force evaluation), K = sum |v|^2 / 2
// with v = y1 / dt (stats/total_energy.cpp:4-10), E = V + K, Rg
// (stats/gyration.cpp:4-18, unit masses), W = 1.5 s_min^2 - 0.5 |s|^2 over
// the singular values s of the centred positions (stats/asphericity.cpp),
// and the native contacts active, |r1 - r2| < thr 2^(-1/6) d_nat on raw
// positions (io/report_structure.cpp:81-85), and their number.
__global__ __launch_bounds__(BLK) void k_observables(Dev D, Derived G, int p, double thr, double *out) {
    __shared__ double lds[BLK / WAVE];
    const int rep = blockIdx.x, tid = threadIdx.x, n = D.nslot;
    const int s0 = D.rep_slot_off[rep], s1 = s0 + D.rep_n[rep];
    const double *P = D.pos[p];
    const double cut = thr * pow(2.0, -1.0 / 6.0);
    double sx = 0, sy = 0, sz = 0, kk = 0, vv = 0, act = 0, tot = 0;
    for (int s = s0 + tid; s < s1; s += BLK) {
        const v3<double> r = ldp(P, s);
        sx += r.x;
        sy += r.y;
        sz += r.z;
        kk += norm_sq(ld3(D.y, n, s));
        vv += D.Vbead[s];
        for (int k = D.nat_start[s]; k < D.nat_start[s + 1]; ++k) {
            const int j = D.nat_j[k];
            if (j <= s) continue;
            tot += 1.0;
            act += norm(ldp(P, j) - r) < cut * D.nat_d[k] ? 1.0 : 0.0;
        }
    }
    const double nb = (double)(s1 - s0);
    const v3<double> c = {block_sum(sx, lds) / nb, block_sum(sy, lds) / nb, block_sum(sz, lds) / nb};
    kk = block_sum(kk, lds);
    vv = block_sum(vv, lds);
    act = block_sum(act, lds);
    tot = block_sum(tot, lds);
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int s = s0 + tid; s < s1; s += BLK) {
        const v3<double> d = ldp(P, s) - c;
        a[0] += d.x * d.x;
        a[1] += d.y * d.y;
        a[2] += d.z * d.z;
        a[3] += d.x * d.y;
        a[4] += d.x * d.z;
        a[5] += d.y * d.z;
    }
    for (int k = 0; k < 6; ++k) a[k] = block_sum(a[k], lds);
    if (tid == 0) {
        double e[3];
        sym3_eig(a, e);
        const double K = 0.5 * kk * G.dt_inv * G.dt_inv;
        double *o = out + 8 * rep;
        o[0] = D.tt[p][rep];
        o[1] = vv;
        o[2] = K;
        o[3] = vv + K;
        o[4] = sqrt((a[0] + a[1] + a[2]) / nb);
        o[5] = 1.5 * e[0] - 0.5 * (e[0] + e[1] + e[2]);
        o[6] = act;
        o[7] = tot;
    }
}

// Per-bead count / fill of the full canonical Verlet list from orig_r
// (export only, legacy_update.cpp:16-62).  mode 0 counts, 1 fills.
__global__ __launch_bounds__(BLK) void k_nl_export(Dev D, Derived G, int *cnt, const int *off, int *oi1, int *oi2,
                                                   uint8_t *onat, int mode) {
    int s = blockIdx.x * BLK + threadIdx.x;
    if (!(D.bflags[s] & B_ACTIVE)) return;
    int rep = D.slot_rep[s];
    Box box = D.box[rep];
    v3<double> rs = ldp(D.orig_r, s);
    int a = D.rep_slot_off[rep], b = a + D.rep_n[rep];
    int k = mode ? off[s] : 0;
    for (int j = s + 1; j < b; ++j) {
        if (D.chain[j] == D.chain[s] && j - s < 3) continue;
        v3<double> d = wrap(box, ldp(D.orig_r, j) - rs);
        if (norm(d) < G.req_r) {
            if (mode) {
                uint8_t kind = 0;  // 0 non_native, 1 native_contacts, 2 native_ssbonds
                for (int q = D.nat_start[s]; q < D.nat_start[s + 1]; ++q)
                    if (D.nat_j[q] == j) kind = 1;
                for (int q = D.ss_start[s]; q < D.ss_start[s + 1]; ++q)
                    if (D.ss_j[q] == j) kind = 2;
                oi1[k] = s;
                oi2[k] = j;
                onat[k] = kind;
            }
            ++k;
        }
    }
    if (!mode) cnt[s] = k;
    (void)a;
}

// rand_gen noise of n consecutive beads of one stream (xmd_rand_normals).
__global__ __launch_bounds__(BLK) void k_rand_normals(Dev D, uint64_t S, long long nb, float *out) {
    long long k = (long long)blockIdx.x * BLK + threadIdx.x;
    int w = (int)(k >> 6);
    uint64_t st = wave_lane_state(D, S, w);
    float nz[3];
    bead_noise(st, NoiseFix{D.bm_fix, D.bm_thr}, nz);
    if (k < nb) {
        out[3 * k] = nz[0];
        out[3 * k + 1] = nz[1];
        out[3 * k + 2] = nz[2];
    }
}

}  // namespace xmd_k
