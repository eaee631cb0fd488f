// lp_reg.h -- ORCA's incremental linear programs over constraints held in
// registers, for the host and gfx950.
//
// A lane's constraint list is two fixed-size arrays, fully unrolled so that
// every constraint stays in a register: OB obstacle slots, of which the
// last nObst hold the obstacle constraints in order (slot i holds constraint
// i - (OB - nObst)), then K agent slots, of which the first na hold the
// agent constraints (slot OB + j holds constraint nObst + j).  Iterating the
// slots in order visits the constraints in the reference's order, so each
// routine performs the reference's operations in the reference's order and
// returns bit-identical results (IEEE binary32 with -ffp-contract=off and
// correctly rounded division and square root, as in ecm_geom.h).
//
// Reference: ORCA::RandomizedLP (ECMAgentSimulator/ORCA.cpp:428-587),
// ORCA::RandomizedLP3D (ORCA.cpp:592-669), Constraint::Contains (ORCA.h:49-60).
#pragma once

#include "ecm_geom.h"

namespace ecm {

// Slot s of the unrolled list (s a compile-time constant after unrolling).
template <int OB, int K>
ECM_HD V2 slot_n(const V2 (&on)[OB], const V2 (&an)[K], int s) { return s < OB ? on[s < OB ? s : 0] : an[s < OB ? 0 : s - OB]; }
template <int OB, int K>
ECM_HD V2 slot_p(const V2 (&op)[OB], const V2 (&ap)[K], int s) { return s < OB ? op[s < OB ? s : 0] : ap[s < OB ? 0 : s - OB]; }

// Reference index of slot s, and the slot of reference index c.
template <int OB>
ECM_HD int slot_to_index(int s, int nObst) { return s < OB ? s - (OB - nObst) : nObst + (s - OB); }
template <int OB>
ECM_HD int index_to_slot(int c, int nObst) { return c < nObst ? OB - nObst + c : OB + (c - nObst); }

// The loop over earlier constraints j < i of RandomizedLP for constraint
// (hn, hp): narrows [lft, rgt]; false where the reference returns i.
template <int OB, int K>
ECM_HD bool lp_narrow(const V2 (&on)[OB], const V2 (&op)[OB], int o_lo, const V2 (&an)[K], const V2 (&ap)[K],
                      uint32_t aval, int i, V2 hn, V2 hp, float& lft, float& rgt) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < OB + K; ++j) {
        if (j < i) {
            const bool vj = j < OB ? (j >= o_lo) : ((aval >> (j < OB ? 0 : j - OB)) & 1u) != 0u;
            if (ok && vj) {
                const V2 jn = slot_n<OB, K>(on, an, j), jp = slot_p<OB, K>(op, ap, j);
                const float den = det(right(hn), right(jn));
                const float num = det(right(jn), hp - jp);
                if (fabsf(den) <= kEps) {
                    if (num < 0.0f) ok = false;
                } else {
                    const float t = num / den;
                    if (den >= 0.0f) rgt = std_min(rgt, t);
                    else lft = std_max(lft, t);
                    if (lft > rgt) ok = false;
                }
            }
        }
    }
    return ok;
}

// RandomizedLP with useDirOpt false over the slots: obstacle slots from o_lo
// on and agent slots whose bit is set in aval.  Returns the slot where the
// reference returns a failing index, or OB + K; out as the reference leaves
// it.
template <int OB, int K>
ECM_HD int lp2_reg(const V2 (&on)[OB], const V2 (&op)[OB], int o_lo, const V2 (&an)[K], const V2 (&ap)[K],
                   uint32_t aval, V2 opt, float maxSpeed, V2& out) {
    if (len(opt) > maxSpeed) out = normalized(opt) * maxSpeed;
    else out = opt;
    int fail = OB + K;
#pragma unroll
    for (int i = 0; i < OB + K; ++i) {
        const bool vi = i < OB ? (i >= o_lo) : ((aval >> (i < OB ? 0 : i - OB)) & 1u) != 0u;
        const V2 hn = slot_n<OB, K>(on, an, i), hp = slot_p<OB, K>(op, ap, i);
        // Constraint::Contains is det(Right(n), p - v) <= 0; a NaN is not contained
        if (fail == OB + K && vi && !(det(right(hn), hp - out) <= 0.0f)) {
            const V2 dir = right(hn);
            const float dpd = dot(dir, hp);
            const float disc = dpd * dpd + maxSpeed * maxSpeed - dot(hp, hp);
            if (disc <= 0.0f) {
                fail = i;
            } else {
                const float sq = sqrtf(disc);
                float lft = -dpd - sq;
                float rgt = -dpd + sq;
                if (!lp_narrow<OB, K>(on, op, o_lo, an, ap, aval, i, hn, hp, lft, rgt)) {
                    fail = i;
                } else {
                    const float t = dot(right(hn), opt - hp);
                    if (t < lft) out = hp + right(hn) * lft;
                    else if (t > rgt) out = hp + right(hn) * rgt;
                    else out = hp + right(hn) * t;
                }
            }
        }
    }
    return fail;
}

// RandomizedLP with useDirOpt true (the form RandomizedLP3D calls) over the
// same kind of list.  False where the reference returns a failing index.
template <int OB, int K>
ECM_HD bool lp_dir_reg(const V2 (&on)[OB], const V2 (&op)[OB], int o_lo, const V2 (&pn)[K], const V2 (&pp)[K],
                       uint32_t pval, V2 opt, float maxSpeed, V2& out) {
    out = opt * maxSpeed;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < OB + K; ++i) {
        const bool vi = i < OB ? (i >= o_lo) : ((pval >> (i < OB ? 0 : i - OB)) & 1u) != 0u;
        const V2 hn = slot_n<OB, K>(on, pn, i), hp = slot_p<OB, K>(op, pp, i);
        if (ok && vi && !(det(right(hn), hp - out) <= 0.0f)) {
            const V2 dir = right(hn);
            const float dpd = dot(dir, hp);
            const float disc = dpd * dpd + maxSpeed * maxSpeed - dot(hp, hp);
            if (disc <= 0.0f) {
                ok = false;
            } else {
                const float sq = sqrtf(disc);
                float lft = -dpd - sq;
                float rgt = -dpd + sq;
                if (!lp_narrow<OB, K>(on, op, o_lo, pn, pp, pval, i, hn, hp, lft, rgt)) ok = false;
                else if (dot(opt, right(hn)) > 0) out = hp + right(hn) * rgt;
                else out = hp + right(hn) * lft;
            }
        }
    }
    return ok;
}

// RandomizedLP3D from reference index `failed` on, out holding the 2D LP's
// result at the failure.  Each lane walks its own candidates ci from its
// failing slot on, skipping those whose penetration is within maxPen (a
// load and a determinant each), and the wavefront runs the bisectors and
// the inner LP once per round for every lane that found its next
// candidate.  A lane needs the inner LP for only one or two candidates, but
// which ones differs from lane to lane: a loop over candidates shared by
// the wavefront would run the union of its lanes' candidates.  The
// bisectors of a round are built for j < jv of each lane (unrolled, no
// register selected by position); fetch(c) returns constraint c as (nx, ny,
// px, py).
template <int OB, int K, typename Fetch>
ECM_HD void lp3d_reg(const V2 (&on)[OB], const V2 (&op)[OB], const V2 (&an)[K], const V2 (&ap)[K], int nObst, int na,
                     int failed, float maxSpeed, Fetch&& fetch, V2& out) {
    const int o_lo = OB - nObst;
    const int end = OB + na;                       // the list's slots are [o_lo, end)
    int v = index_to_slot<OB>(failed, nObst);
    float maxPen = 0.0f;
    while (true) {
        // this lane's next candidate whose penetration exceeds maxPen
        V2 cn = mk(0.0f, 0.0f), cp = mk(0.0f, 0.0f);
        bool found = false;
        for (; v < end; ++v) {
            const float4 c4 = fetch(slot_to_index<OB>(v, nObst));
            cn = mk(c4.x, c4.y);
            cp = mk(c4.z, c4.w);
            if (!(det(right(cn), cp - out) <= maxPen)) {
                found = true;
                break;
            }
        }
        if (!found) break;
        const int jv = v - OB;
        const V2 dir = right(cn);
        // bisectors with the agent constraints before ci (ORCA.cpp:614-646)
        V2 pn[K], pp[K];
        uint32_t pval = 0u;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            pn[j] = mk(0.0f, 0.0f);
            pp[j] = mk(0.0f, 0.0f);
            if (j < jv) {
                const float d = det(right(cn), right(an[j]));
                if (fabsf(d) <= kEps) {
                    if (!(dot(cn, an[j]) > 0)) {
                        pp[j] = (cp + ap[j]) * 0.5f;
                        pn[j] = normalized(an[j] - cn);
                        pval |= 1u << j;
                    }
                } else {
                    const float tt = det(right(an[j]), cp - ap[j]) / d;
                    pp[j] = cp + right(cn) * tt;
                    pn[j] = normalized(an[j] - cn);
                    pval |= 1u << j;
                }
            }
        }
        const V2 temp = out;
        if (!lp_dir_reg<OB, K>(on, op, o_lo, pn, pp, pval, cn, maxSpeed, out)) out = temp;
        maxPen = det(dir, cp - out);
        ++v;
    }
}

}  // namespace ecm
