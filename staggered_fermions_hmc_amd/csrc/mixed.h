// Two-kernel mixed-precision CG iteration (cgmixed.hip).
#pragma once

#include "hopcore.h"
#include "internal.h"

namespace sfh {

// Sloppy vectors of one mixed-precision solve.  In the fused form r and p are
// double-buffered: the iteration at position j of a chunk (chunks have an
// even length) writes r_k and p_k into buffer j & 1 while other workgroups
// still read r_{k-1} and p_{k-1} from the other buffer; a chunk starts from
// buffer 1.  In the split form both slots point to one array.  Norms are the per-site
// max-norms of FP16 storage (unused in FP32).
template <class T> struct MixBufs {
  typename Prec<T>::S *r[2], *p[2], *v, *t;
  float *rn[2], *pn[2], *vn, *tn;
  float2* x;  // FP32 solution accumulated since the last reliable update
};

// Kernel A of iteration k: r_k = r_{k-1} - alpha v_{k-1}, p_k = r_k +
// beta p_{k-1} and x += alpha p_{k-1} at each site, with r_k and p_k
// recomputed at the neighbours, then t = D^dag p_k.  With st_in != st_out
// (small grids, see mix_defers) it first finishes the previous B's dots from
// `partials` and writes the advanced state to st_out.
// r12: SU(3) links, read as two rows with the third rebuilt.  cur: the
// buffer written (the iteration's position in the chunk, mod 2).
template <class T>
void mix_a(hipStream_t s, const Geom& g, const typename Prec<T>::L* U, const MixBufs<T>& b, float mass,
           const CGState* st_in, CGState* st_out, const double* partials, bool r12, int cur);
// How kernel B's four dots reach the finisher: a grid reduction in B with
// its own finisher (kDotsFinish); plain per-workgroup partials that every
// workgroup of the next fused kernel A sums (kDotsPartials: small grids);
// group sums of the grid reduction's first level that every workgroup of the
// next split-form update kernel adds up (kDotsGroups).  With either deferred
// form, iteration j of a chunk reads state copy j & 1 and writes copy
// (j + 1) & 1.
constexpr int kDotsFinish = 0, kDotsPartials = 1, kDotsGroups = 2;
// Whether the fused form's B can leave per-workgroup partials: true when the
// grid is small enough for every workgroup of A to read them all.
bool mix_defers(const Geom& g);
// Split form of kernel A (for lattices whose neighbour vectors do not stay
// in cache): the own-site update alone, in place; the caller then runs the
// plain hop t = D^dag p_k.  It first finishes the previous B's dots from
// their group sums (kDotsGroups), reading the state from st_in and writing
// the advanced state to st_out.
template <class T>
void mix_upd(hipStream_t s, const Geom& g, const MixBufs<T>& b, const CGState* st_in, CGState* st_out,
             const double* partials);
// Fused form chosen below this many sites; the split form at and above.
constexpr int kMixFusedMaxV = 1 << 18;
// Kernel B of iteration k: v = D t and the four dots of FinFused, with the
// r_k and p_k of buffer cur (the same as the A before it).
template <class T>
void mix_b(hipStream_t s, const Geom& g, const typename Prec<T>::L* U, const MixBufs<T>& b, float mass, CGState* st,
           double* partials, int dm, bool r12, int cur);
// Reliable update, first and last steps (gated on the solve state):
// x64 += x + alpha p_k (p_k from buffer st->pbuf) and x = 0; then r = r64 in
// sloppy storage, into buffer 1 with p_k moved there, where the next chunk
// starts.  With deferred dots (dm not kDotsFinish) mix_acc first finishes
// the chunk's last B from `partials`, reading the state from st_in and
// writing it to st_out, the copy the rest of the update uses; mix_apply
// then copies that state back to st_out, where the next chunk starts.
template <class T>
void mix_acc(hipStream_t s, const Geom& g, double2* x64, const MixBufs<T>& b, const CGState* st_in, CGState* st_out,
             const double* partials, int dm);
// Reliable update, residual: r = phi - D t with the FP64 links (t = D^dag x
// from the first FP64 hop) and |r|^2 to the update's finisher, in one
// launch gated on the pending update.  r12: the links pass the FP64 SU(3)
// check (two rows read).
void ru_resid(hipStream_t s, const Geom& g, const double2* U, const double2* t, const double2* phi, double2* r,
              double mass, CGState* st, double* partials, bool r12);
template <class T>
void mix_apply(hipStream_t s, const Geom& g, const double2* r64, const MixBufs<T>& b, const CGState* st,
               CGState* st_out);

}  // namespace sfh
