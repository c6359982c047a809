// Host-side G2 arithmetic for the tail of a G2 MSM: the window fold and the
// XYZZ -> Jacobian conversion over Fp2 = Fp[i] / (i^2 + 1), on the Fp
// functions of host_ec.h.  Layouts match the device's xyzz2 (X, Y, ZZ, ZZZ,
// each (c0, c1)) and blst_p2 (X, Y, Z).
#pragma once
#include "host_ec.h"

namespace hostec2 {

using hostec::fp6;

struct f2 { fp6 c0, c1; };
struct xyzz12 { f2 x, y, zz, zzz; };
struct jac12 { f2 x, y, z; };

inline bool is_zero(const f2& a) { return hostec::is_zero(a.c0) && hostec::is_zero(a.c1); }
inline void add(f2& r, const f2& a, const f2& b) {
    hostec::add(r.c0, a.c0, b.c0);
    hostec::add(r.c1, a.c1, b.c1);
}
inline void sub(f2& r, const f2& a, const f2& b) {
    hostec::sub(r.c0, a.c0, b.c0);
    hostec::sub(r.c1, a.c1, b.c1);
}
inline void mul(f2& r, const f2& a, const f2& b) {
    fp6 t0, t1, t2, t3;
    hostec::mul(t0, a.c0, b.c0);
    hostec::mul(t1, a.c1, b.c1);
    hostec::mul(t2, a.c0, b.c1);
    hostec::mul(t3, a.c1, b.c0);
    hostec::sub(r.c0, t0, t1);
    hostec::add(r.c1, t2, t3);
}

inline bool is_inf(const xyzz12& p) { return is_zero(p.zz); }

// dbl-2008-s-1
inline void dbl(xyzz12& r, const xyzz12& p) {
    if (is_inf(p)) { r = p; return; }
    f2 U, V, W, S, M, t;
    add(U, p.y, p.y);
    mul(V, U, U);
    mul(W, U, V);
    mul(S, p.x, V);
    mul(M, p.x, p.x);
    add(t, M, M);
    add(M, M, t);
    xyzz12 o;
    mul(o.x, M, M);
    add(t, S, S);
    sub(o.x, o.x, t);
    sub(t, S, o.x);
    mul(t, t, M);
    mul(o.y, W, p.y);
    sub(o.y, t, o.y);
    mul(o.zz, V, p.zz);
    mul(o.zzz, W, p.zzz);
    r = o;
}

// add-2008-s with the equal and opposite cases resolved
inline void addp(xyzz12& r, const xyzz12& p, const xyzz12& q) {
    if (is_inf(q)) { r = p; return; }
    if (is_inf(p)) { r = q; return; }
    f2 U1, S1, P, R;
    mul(U1, p.x, q.zz);
    mul(S1, p.y, q.zzz);
    mul(P, q.x, p.zz);
    mul(R, q.y, p.zzz);
    sub(P, P, U1);
    sub(R, R, S1);
    if (is_zero(P)) {
        if (is_zero(R)) dbl(r, p);
        else memset(&r, 0, sizeof(r));
        return;
    }
    f2 PP, PPP, Q, t;
    xyzz12 o;
    mul(PP, P, P);
    mul(PPP, PP, P);
    mul(Q, U1, PP);
    mul(t, R, R);
    sub(t, t, PPP);
    sub(t, t, Q);
    sub(o.x, t, Q);
    sub(Q, Q, o.x);
    mul(Q, Q, R);
    mul(t, S1, PPP);
    sub(o.y, Q, t);
    mul(t, p.zz, q.zz);
    mul(o.zz, t, PP);
    mul(t, p.zzz, q.zzz);
    mul(o.zzz, t, PPP);
    r = o;
}

// Pointxyzz::to_Jacobian; infinity maps to the all-zero blst_p2
inline void to_jac(jac12& r, const xyzz12& p) {
    if (is_inf(p)) { memset(&r, 0, sizeof(r)); return; }
    mul(r.x, p.x, p.zz);
    mul(r.y, p.y, p.zzz);
    r.z = p.zz;
}

// sum_j 2^(c j) w[j], Horner from the top window
inline void fold_windows(xyzz12& out, const xyzz12* w, int nwin, int c) {
    xyzz12 acc;
    memset(&acc, 0, sizeof(acc));
    for (int j = nwin - 1; j >= 0; j--) {
        if (j != nwin - 1)
            for (int k = 0; k < c; k++) dbl(acc, acc);
        addp(acc, acc, w[j]);
    }
    out = acc;
}

}  // namespace hostec2
