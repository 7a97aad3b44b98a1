// Replays the reference's known-answer tests for the hot path through the
// C++ mirror (include/la_matrix.hpp), i.e. through the C ABI on the GPU.
// Each case names the reference test it restates.  Exit status 0 = all pass.
#include "la_matrix.hpp"

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

using la::Matrix;
using la::m;

static int failures = 0;

static void run(const char *name, const std::function<bool()> &f) {
    bool ok = false;
    try {
        ok = f();
    } catch (const std::exception &e) {
        std::printf("EXC  %s: %s\n", name, e.what());
    }
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
    if (!ok) ++failures;
}

template <typename F>
static bool panics(F f) {
    try {
        f();
    } catch (const la::Panic &) {
        return true;
    }
    return false;
}

static bool approx_eq(const Matrix<double> &a, const Matrix<double> &b) {
    // ApproxEq for f64: |a - b| < 1e-6 elementwise (src/approxeq.rs:45)
    if (a.rows() != b.rows() || a.cols() != b.cols()) return false;
    for (size_t i = 0; i < a.get_data().size(); ++i)
        if (!(std::fabs(a.get_data()[i] - b.get_data()[i]) < 1e-6)) return false;
    return true;
}

int main() {
    using D = double;
    // ---------------------------------------------------------- lu.rs
    run("lu.rs:281 test_lu_square", [] {
        auto a = m<D>({{1, 2, 0}, {3, 6, -1}, {1, 2, 1}});
        la::LUDecomposition<D> lu(a);
        return lu.get_l() * lu.get_u() == lu.get_p() * a;
    });
    run("lu.rs:291 test_lu2_m_over_n", [] {
        auto a = m<D>({{1, 2}, {3, 4}, {5, 6}});
        la::LUDecomposition<D> lu(a);
        return lu.get_l() * lu.get_u() == lu.get_p() * a;
    });
    run("lu.rs:301 test_lu2_m_under_n", [] {
        auto a = m<D>({{1, 2, 3}, {4, 5, 6}});
        la::LUDecomposition<D> lu(a);
        return lu.get_l() * lu.get_u() == lu.get_p() * a;
    });
    run("lu.rs:311 lu_solve_test", [] {
        la::LUDecomposition<D> lu(m<D>({{2, 1, 0}, {1, 1, 0}, {0, 0, 1}}));
        return approx_eq(*lu.solve(m<D>({{1}, {2}, {3}})), m<D>({{-1}, {3}, {3}}));
    });
    run("lu.rs:319 lu_solve_test_incompatible (should_panic)", [] {
        la::LUDecomposition<D> lu(m<D>({{2, 1, 0}, {1, 1, 0}, {0, 0, 1}}));
        return panics([&] { (void)lu.solve(m<D>({{1}, {2}, {3}, {4}})); });
    });
    run("lu.rs:328 lu_solve_test_singular", [] {
        la::LUDecomposition<D> lu(m<D>({{2, 6}, {1, 3}}));
        return !lu.solve(m<D>({{1}, {2}})).has_value();
    });
    run("lu.rs:336 lu_is_singular_test", [] {
        return la::LUDecomposition<D>(m<D>({{2, 6}, {1, 3}})).is_singular() &&
               !la::LUDecomposition<D>(m<D>({{2, 6}, {1, 4}})).is_singular();
    });
    run("lu.rs:347 lu_is_non_singular_test", [] {
        return la::LUDecomposition<D>(m<D>({{4, 8}, {3, 4}})).is_non_singular() &&
               !la::LUDecomposition<D>(m<D>({{4, 6}, {2, 3}})).is_non_singular();
    });
    run("lu.rs:358 lu_det_test", [] {
        return la::LUDecomposition<D>(m<D>({{4, 8}, {3, 4}})).det() == -8.0 &&
               la::LUDecomposition<D>(m<D>({{4, 8}, {2, 4}})).det() == 0.0;
    });
    run("lu.rs:369 lu_det_test_not_square (should_panic)", [] {
        la::LUDecomposition<D> lu(m<D>({{1, 2, 3}, {4, 5, 6}}));
        return panics([&] { (void)lu.det(); });
    });
    // ---------------------------------------------------- cholesky.rs
    run("cholesky.rs:147 cholesky_square_pos_def_test", [] {
        auto a = m<D>({{4, 12, -16}, {12, 37, -43}, {-16, -43, 98}});
        auto c = la::CholeskyDecomposition<D>::create(a);
        return c && c->get_l() * c->get_l().t() == a &&
               c->get_l().get_data() == std::vector<D>{2, 0, 0, 6, 1, 0, -8, 5, 3};
    });
    run("cholesky.rs:155 cholesky_not_pos_def_test", [] {
        return !la::CholeskyDecomposition<D>::create(m<D>({{4, 12, -16}, {12, 37, 43}, {-16, 43, 98}}));
    });
    run("cholesky.rs:161 cholesky_not_square_test", [] {
        return !la::CholeskyDecomposition<D>::create(m<D>({{4, 12, -16}, {12, 37, 43}}));
    });
    run("cholesky.rs:167 cholesky_solve_test", [] {
        auto c = la::CholeskyDecomposition<D>::create(m<D>({{2, 1, 0}, {1, 1, 0}, {0, 0, 1}}));
        return c && approx_eq(c->solve(m<D>({{1}, {2}, {3}})), m<D>({{-1}, {3}, {3}}));
    });
    run("cholesky.rs:175 cholesky_solve_test_incompatible (should_panic)", [] {
        auto c = la::CholeskyDecomposition<D>::create(m<D>({{2, 1, 0}, {1, 1, 0}, {0, 0, 1}}));
        return panics([&] { (void)c->solve(m<D>({{1}, {2}, {3}, {4}})); });
    });
    // ---------------------------------------------------------- mod.rs
    run("mod.rs:1479 test_mul (integer, reference loop)", [] {
        return (m<long>({{1, 2}, {3, 4}}) * m<long>({{3, 4}, {5, 6}})).get_data() ==
               std::vector<long>{13, 16, 29, 36};
    });
    run("mod.rs:1479 test_mul (f64, GPU)", [] {
        return (m<D>({{1, 2}, {3, 4}}) * m<D>({{3, 4}, {5, 6}})).get_data() ==
               std::vector<D>{13, 16, 29, 36};
    });
    run("mod.rs:1486 test_mul_incompatible (should_panic)", [] {
        return panics([] { (void)(m<D>({{1, 2}, {3, 4}}) * m<D>({{1, 2}, {3, 4}, {5, 6}})); });
    });
    run("mmatrix.rs:234 test_mmul", [] {
        auto dst = Matrix<D>::dirty(2, 2);
        m<D>({{1, 2}, {3, 4}}).mmul(m<D>({{3, 4}, {5, 6}}), dst);
        return dst.get_data() == std::vector<D>{13, 16, 29, 36};
    });
    run("mmatrix.rs:243 test_mmul_incompatible_src (should_panic)", [] {
        auto dst = Matrix<D>::dirty(2, 2);
        return panics([&] { m<D>({{1, 2}, {3, 4}}).mmul(m<D>({{1, 2}, {3, 4}, {5, 6}}), dst); });
    });
    run("mmatrix.rs:252 test_mmul_incompatible_dst (should_panic)", [] {
        auto dst = Matrix<D>::dirty(3, 3);
        return panics([&] { m<D>({{1, 2}, {3, 4}}).mmul(m<D>({{1, 2}, {3, 4}}), dst); });
    });
    run("mod.rs:1506 test_det", [] {
        return std::fabs(m<D>({{6, -7, 10}, {0, 3, -1}, {0, 5, -7}}).det() - -96.0) <= 1.1920929e-7;
    });
    run("mod.rs:1512 test_det_not_square (should_panic)", [] {
        return panics([] { (void)m<D>({{6, -7, 10}, {0, 3, -1}}).det(); });
    });
    run("mod.rs:1518 test_solve (exact)", [] {
        auto x = m<D>({{1, 1, 1}, {1, -1, 4}, {2, 3, -5}}).solve(m<D>({{3}, {4}, {0}}));
        return x && *x == m<D>({{1}, {1}, {1}});
    });
    run("mod.rs:1527 test_inverse", [] {
        std::vector<D> d{16, -1, 23, 0, 42, -6, 0, 30, -18};
        for (auto &v : d) v /= 96.0;
        auto inv = m<D>({{6, -7, 10}, {0, 3, -1}, {0, 5, -7}}).inverse();
        return inv && approx_eq(*inv, Matrix<D>(3, 3, d));
    });
    run("mod.rs:1535 test_inverse_not_square (should_panic)", [] {
        return panics([] { (void)m<D>({{6, -7, 10}, {0, 3, -1}}).inverse(); });
    });
    run("mod.rs:1542 test_inverse_singular", [] {
        return !m<D>({{2, 6}, {1, 3}}).inverse().has_value();
    });
    run("mod.rs:1554 test_is_singular", [] { return m<D>({{2, 6}, {1, 3}}).is_singular(); });
    run("mod.rs:1560 test_is_singular_non_square (should_panic)", [] {
        return panics([] { (void)m<D>({{1, 2, 3}, {4, 5, 6}}).is_singular(); });
    });
    run("mod.rs:1567 test_is_non_singular", [] { return m<D>({{2, 6}, {6, 3}}).is_non_singular(); });
    run("qr.rs:241 qr_test", [] {
        auto a = m<D>({{12, -51, 4}, {6, 167, -68}, {-4, 24, -41}});
        la::QRDecomposition<D> qr(a);
        return approx_eq(qr.get_q() * qr.get_r(), a);
    });
    run("qr.rs:248 qr_test_m_over_n", [] {
        auto a = m<D>({{1, 2}, {3, 4}, {5, 6}});
        la::QRDecomposition<D> qr(a);
        return approx_eq(qr.get_q() * qr.get_r(), a);
    });
    run("qr.rs:255 qr_test_n_over_m", [] {
        auto a = m<D>({{1, 2, 3}, {4, 5, 6}});
        la::QRDecomposition<D> qr(a);
        return approx_eq(qr.get_q() * qr.get_r(), a);
    });
    run("qr.rs:196 solve (least squares, square)", [] {
        auto a = m<D>({{12, -51, 4}, {6, 167, -68}, {-4, 24, -41}});
        auto x = la::QRDecomposition<D>(a).solve(m<D>({{1}, {2}, {3}}));
        return x && approx_eq(a * *x, m<D>({{1}, {2}, {3}}));
    });
    run("qr.rs:108 is_full_rank panics when cols > rows", [] {
        return panics([] { (void)la::QRDecomposition<D>(m<D>({{1, 2, 3}, {4, 5, 6}})).is_full_rank(); });
    });
    run("mod.rs:1549 test_pinverse", [] {
        auto a = m<D>({{1, 2}, {3, 4}, {5, 6}});
        return approx_eq(a.pinverse() * a, Matrix<D>::id(2, 2));
    });
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
