// Host check of the register transforms' lazy arithmetic (montmul_amd/csrc/mm_v2.h): every
// operation is compared with exact residues (128-bit products and %) on boundary and random
// inputs at the top of each lazy range, and the radix-8/4 butterflies with a direct DFT over
// omega_8 = 2^h (1 + i). Test infrastructure: built and run by tests/test_arith.py.
#include <cstdio>
#include <cstdlib>

#include "../../montmul_amd/csrc/mm_v2.h"

using namespace mm;
using namespace mm::v2;

static int g_fail = 0;
static long g_checks = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        g_checks++;                                     \
        if (!(c)) {                                     \
            if (g_fail++ < 20) {                        \
                std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
                std::printf(__VA_ARGS__);               \
                std::printf("\n");                      \
            }                                           \
        }                                               \
    } while (0)

static uint64_t rs = 0x9e3779b97f4a7c15ull;
static uint64_t rnd() {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return rs;
}

// Exact complex product mod P2 and mod P1 (canonical).
static F2 ex2(uint64_t ar, uint64_t ai, uint64_t br, uint64_t bi) {
    u128 p = P2;
    u128 re = ((u128)(ar % P2) * (br % P2) + (u128)(ai % P2) * ((P2 - bi % P2) % P2)) % p;
    u128 im = ((u128)(ar % P2) * (bi % P2) + (u128)(ai % P2) * (br % P2)) % p;
    return F2{(uint64_t)re, (uint64_t)im};
}
static F1 ex1(uint64_t ar, uint64_t ai, uint64_t br, uint64_t bi) {
    uint64_t re = ((ar % P1) * (br % P1) + (ai % P1) * ((P1 - bi % P1) % P1)) % P1;
    uint64_t im = ((ar % P1) * (bi % P1) + (ai % P1) * (br % P1)) % P1;
    return F1{(uint32_t)re, (uint32_t)im};
}

static const uint64_t LAZY2 = P2 + 7;  // lazy M61 words are < 2^61 + 7
static uint64_t lazy2() {
    static const uint64_t edge[] = {0, 1, 2, P2 - 1, P2, P2 + 1, P2 + 6, 1ull << 60, (1ull << 61) - 2, 0x7fffffffull, 0x80000000ull};
    uint64_t r = rnd();
    if ((r & 3) == 0) return edge[(r >> 2) % (sizeof(edge) / sizeof(edge[0]))];
    return rnd() % LAZY2;
}
static uint64_t canon2r() {
    static const uint64_t edge[] = {0, 1, P2 - 1, 0x7fffffffull, 0x80000000ull, (1ull << 61) - 0x80000000ull};
    uint64_t r = rnd();
    if ((r & 3) == 0) return edge[(r >> 2) % (sizeof(edge) / sizeof(edge[0]))];
    return rnd() % P2;
}
static uint32_t lazy1() {
    static const uint32_t edge[] = {0, 1, P1 - 1, P1, 1u << 30};
    uint64_t r = rnd();
    if ((r & 3) == 0) return edge[(r >> 2) % 5];
    return (uint32_t)(rnd() % (P1 + 1ull));
}
static bool ok2(uint64_t v, uint64_t want) { return v < LAZY2 && v % P2 == want % P2; }
static bool ok1(uint32_t v, uint64_t want) { return v <= P1 && v % P1 == want % P1; }

static void check_m61_ops(int iters) {
    for (int it = 0; it < iters; it++) {
        uint64_t a = lazy2(), b = lazy2();
        CHECK(ok2(ad2(a, b), (a % P2 + b % P2)), "ad2 %llx %llx", (unsigned long long)a, (unsigned long long)b);
        CHECK(ok2(sb2(a, b), (a % P2 + P2 - b % P2)), "sb2");
        CHECK(ok2(rot2(a), (uint64_t)(((u128)(a % P2) << 30) % P2)), "rot2");
        CHECK(canon2(a) == a % P2, "canon2");
        Z2 x{lazy2(), lazy2()};
        F2 c{canon2r(), canon2r()};
        F2 w = ex2(x.re, x.im, c.re, c.im);
        Z2 r = Op2::mul(x, spl_c(c));
        CHECK(ok2(r.re, w.re) && ok2(r.im, w.im), "mul x=(%llx,%llx) c=(%llx,%llx)", (unsigned long long)x.re,
              (unsigned long long)x.im, (unsigned long long)c.re, (unsigned long long)c.im);
        Z2 r2 = Op2::mul(x, c);
        CHECK(r2.re == r.re && r2.im == r.im, "mul F2");
        F2 q = ex2(x.re, x.im, x.re, x.im);
        Z2 s = Op2::sqr(x);
        CHECK(ok2(s.re, q.re) && ok2(s.im, q.im), "sqr x=(%llx,%llx)", (unsigned long long)x.re, (unsigned long long)x.im);
        uint32_t cs = 1 + (uint32_t)(rnd() % 4);
        Z2 sc = Op2::scale(x, cs);
        CHECK(ok2(sc.re, (uint64_t)((u128)(x.re % P2) * cs % P2)) && ok2(sc.im, (uint64_t)((u128)(x.im % P2) * cs % P2)), "scale");
        uint32_t v = (it & 7) == 0 ? 0xffffffffu : (uint32_t)rnd();
        Z2 rm = Op2::real_mul(v, spl_c(c));
        F2 wm = ex2(v, 0, c.re, c.im);
        CHECK(ok2(rm.re, wm.re) && ok2(rm.im, wm.im), "real_mul");
        Z2 cj = Op2::conj(x);
        CHECK(ok2(cj.re, x.re) && ok2(cj.im, (P2 - x.im % P2)), "conj");
    }
}

static void check_m31_ops(int iters) {
    for (int it = 0; it < iters; it++) {
        uint32_t a = lazy1(), b = lazy1();
        CHECK(ok1(ad1(a, b), (uint64_t)a + b), "ad1");
        CHECK(ok1(sb1(a, b), (uint64_t)a + P1 - b % P1), "sb1");
        CHECK(ok1(rot1(a), ((uint64_t)(a % P1) << 15) % P1), "rot1");
        Z1 x{lazy1(), lazy1()};
        F1 c{(uint32_t)(rnd() % P1), (uint32_t)(rnd() % P1)};
        F1 w = ex1(x.re, x.im, c.re, c.im);
        Z1 r = Op1::mul(x, c);
        CHECK(ok1(r.re, w.re) && ok1(r.im, w.im), "mul1");
        F1 q = ex1(x.re, x.im, x.re, x.im);
        Z1 s = Op1::sqr(x);
        CHECK(ok1(s.re, q.re) && ok1(s.im, q.im), "sqr1");
        uint32_t v = (uint32_t)rnd();
        Z1 rm = Op1::real_mul(v, c);
        F1 wm = ex1(v, 0, c.re, c.im);
        CHECK(ok1(rm.re, wm.re) && ok1(rm.im, wm.im), "real_mul1");
    }
}

static uint32_t rev3(uint32_t m) { return ((m & 1) << 2) | (m & 2) | (m >> 2); }

// omega_8 = 2^h (1 + i) and its powers, canonical.
template <class O> struct Ref;
template <> struct Ref<Op2> {
    static F2 w(int j) {
        F2 r{1, 0}, b{1ull << 30, 1ull << 30};
        for (int q = 0; q < ((j % 8) + 8) % 8; q++) r = ex2(r.re, r.im, b.re, b.im);
        return r;
    }
    static F2 mul(F2 a, F2 b) { return ex2(a.re, a.im, b.re, b.im); }
    static F2 add(F2 a, F2 b) { return F2{(a.re + b.re) % P2, (a.im + b.im) % P2}; }
    static F2 of(Z2 z) { return F2{z.re % P2, z.im % P2}; }
    static Z2 in() { return Z2{lazy2(), lazy2()}; }
    static bool lazy(Z2 z) { return z.re < LAZY2 && z.im < LAZY2; }
};
template <> struct Ref<Op1> {
    static F1 w(int j) {
        F1 r{1, 0}, b{1u << 15, 1u << 15};
        for (int q = 0; q < ((j % 8) + 8) % 8; q++) r = ex1(r.re, r.im, b.re, b.im);
        return r;
    }
    static F1 mul(F1 a, F1 b) { return ex1(a.re, a.im, b.re, b.im); }
    static F1 add(F1 a, F1 b) { return F1{(uint32_t)(((uint64_t)a.re + b.re) % P1), (uint32_t)(((uint64_t)a.im + b.im) % P1)}; }
    static F1 of(Z1 z) { return F1{z.re % P1, z.im % P1}; }
    static Z1 in() { return Z1{lazy1(), lazy1()}; }
    static bool lazy(Z1 z) { return z.re <= P1 && z.im <= P1; }
};

// dif8: slot m holds X_rev3(m), X_f = sum_j v_j omega_8^(j f); dit8 inverts it up to the factor 8
// (inverse roots, bit-reversed in, natural out); dif4/dit4 likewise over omega_4 = i.
template <class O> static void check_bfly(int iters, const char* name) {
    typedef Ref<O> R;
    typedef typename O::Z Z;
    for (int it = 0; it < iters; it++) {
        Z v[8], u[8];
        for (int j = 0; j < 8; j++) v[j] = u[j] = R::in();
        dif8<O>(u);
        for (int m = 0; m < 8; m++) {
            uint32_t f = rev3(m);
            auto want = R::of(Z{});
            for (int j = 0; j < 8; j++) want = R::add(want, R::mul(R::of(v[j]), R::w((int)(j * f))));
            auto got = R::of(u[m]);
            CHECK(R::lazy(u[m]) && got.re == want.re && got.im == want.im, "%s dif8 slot %d", name, m);
        }
        Z y[8];
        for (int m = 0; m < 8; m++) y[m] = R::in();
        Z z[8];
        for (int m = 0; m < 8; m++) z[m] = y[m];
        dit8<O>(z);
        for (int j = 0; j < 8; j++) {
            auto want = R::of(Z{});
            for (int m = 0; m < 8; m++) want = R::add(want, R::mul(R::of(y[m]), R::w(-(int)(j * rev3(m)))));
            auto got = R::of(z[j]);
            CHECK(R::lazy(z[j]) && got.re == want.re && got.im == want.im, "%s dit8 out %d", name, j);
        }
        // radix 4 halves: omega_4 = omega_8^2
        Z a[8], b[8];
        for (int m = 0; m < 8; m++) a[m] = b[m] = R::in();
        dif4<O>(b);
        dif4<O>(b + 4);
        for (int g = 0; g < 2; g++)
            for (int m = 0; m < 4; m++) {
                uint32_t f = ((m & 1) << 1) | (m >> 1);
                auto want = R::of(Z{});
                for (int j = 0; j < 4; j++) want = R::add(want, R::mul(R::of(a[4 * g + j]), R::w((int)(2 * j * f))));
                auto got = R::of(b[4 * g + m]);
                CHECK(R::lazy(b[4 * g + m]) && got.re == want.re && got.im == want.im, "%s dif4", name);
            }
        for (int m = 0; m < 8; m++) b[m] = a[m];
        dit4<O>(b);
        for (int j = 0; j < 4; j++) {
            auto want = R::of(Z{});
            for (int m = 0; m < 4; m++) want = R::add(want, R::mul(R::of(a[m]), R::w(-(int)(2 * j * (((m & 1) << 1) | (m >> 1))))));
            auto got = R::of(b[j]);
            CHECK(R::lazy(b[j]) && got.re == want.re && got.im == want.im, "%s dit4", name);
        }
    }
}

int main(int argc, char** argv) {
    int iters = argc > 1 ? std::atoi(argv[1]) : 200000;
    check_m61_ops(iters);
    check_m31_ops(iters);
    check_bfly<Op2>(iters / 20, "M61");
    check_bfly<Op1>(iters / 20, "M31");
    std::printf("%s: %ld checks, %d failures\n", g_fail ? "FAIL" : "OK", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
