// pb_oracle.cpp - FP64 CPU restatement of PlaBundle's online point-Schur frame
// bundle adjustment. TEST INFRASTRUCTURE ONLY: the parity checker and the CPU
// baseline. Nothing in the shipped library links, loads or calls this file.
//
// It follows (paths relative to the PlaBundle tree):
//   controller ............ src/internal/BundleAdjustPlaMatrix.cpp:77-158, 162-631 (online Schur branch)
//   active problem ........ src/internal/BundleAdjustPlaMatrixProblem.cpp:29-278
//   gauge normalisation ... src/internal/BundleAdjustValidation.cpp:261-326, 371-896 (image-only subset)
//   intrinsic groups ...... src/internal/BundleAdjustPlaMatrixModel.cpp:84-269
//   projection adapter .... src/internal/BundleAdjustPlaMatrixProjection.cpp:26-285, 319-354
//   camera parameters ..... src/internal/CameraState.cpp:164-204, 387-481
//   primary terms ......... src/internal/BundleAdjustPlaMatrixAssembly.cpp:59-128, 186-342, 442-530
//   retraction ............ src/internal/BundleAdjustPlaMatrixAssembly.cpp:101-179, 1111-1247
//   online Schur .......... src/internal/BundleAdjustPlaMatrixReferenceSchur.cpp:140-239, 262-536, 675-1082
//
// The frame camera model itself lives in PlaCamera, which is not part of the
// reference tree. It is restated here from the in-tree closed form
// (src/internal/BundleAdjustProjection.h:58-91, OpenCV Brown) and the public
// Agisoft Metashape manual (Appendix "Camera models", frame cameras) for the
// Metashape convention. Choices for parts only PlaMatrix/PlaCamera define
// (damping form, zero-diagonal rule, 3x3 inverse tolerance, median rule) are
// marked "unpinned choice" below.
//
// Parity pins (tests/test_oracle_*.py): central differences of every
// Jacobian, robust-loss closed forms, the reference's own pinhole known answer
// (test/test_native_problem.cpp:38-49), dense full-system vs Schur-reduced
// solves, hand-built block-numbering fixtures and controller invariants.

#include "../include/pb_hip.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
    constexpr int kP = 13;  // primary block size (BundleAdjustPlaMatrixProblem.h:41)
    constexpr int kE = 3;   // eliminated block size (BundleAdjustPlaMatrixProblem.h:42)
    constexpr double kInf = std::numeric_limits<double>::infinity();

    using Mat3 = std::array<double, 9>;
    using Vec3 = std::array<double, 3>;
    using Slots = std::array<double, kP>;
    using SlotMask = std::array<bool, kP>;

    struct InvalidProjection : std::runtime_error
    {
        InvalidProjection() : std::runtime_error("invalid projection") {}
    };

    struct Camera
    {
        Mat3 R{};  // camera-to-world, row-major
        Vec3 C{};
        Slots p{};  // [fx, log(fy/fx), cx, cy, k1, k2, k3, p1, p2, b2, k4, p3, p4]
        int conv = PB_BROWN_OPENCV;
        int width = 0;
        int height = 0;
    };

    struct Observation
    {
        int cam = -1;
        double u = 0.0, v = 0.0, weight = 1.0, scale = 1.0;
    };

    // Survey constraints of one track (constraints.h:39-46, 73-88). Control points keep their whitening
    // matrix A (residual sqrt(w) A (X - P)); laser planes their plane point and (unnormalised) normal.
    struct ControlPoint
    {
        int index = -1;  // position in the caller's control-point list
        Vec3 P{};
        Mat3 A{};
        double weight = 1.0;
    };

    struct LaserPlane
    {
        Vec3 p{};
        Vec3 n{};
        double weight = 1.0;
    };

    struct Track
    {
        Vec3 X{};
        std::vector<Observation> obs;
        std::vector<LaserPlane> planes;
        std::vector<ControlPoint> controls;
    };

    // makeControlPointWhitening (src/control_point.cpp:91-165): isotropic -> diag(1/sigma); covariance S = L L^T
    // -> A = L^-1; square-root information -> A as given, with A^T A required positive definite. The Cholesky
    // test of the reference (every pivot L_ii^2 above a threshold) is restated with threshold 0.
    bool cholesky3(const Mat3& m, Mat3* L)
    {
        for (double v : m)
            if (!std::isfinite(v)) return false;
        *L = Mat3{};
        for (int i = 0; i < 3; ++i)
        {
            for (int j = 0; j <= i; ++j)
            {
                double s = 0.5 * (m[i * 3 + j] + m[j * 3 + i]);
                for (int k = 0; k < j; ++k) s -= (*L)[i * 3 + k] * (*L)[j * 3 + k];
                if (i == j)
                {
                    if (!(s > 0.0) || !std::isfinite(s)) return false;
                    (*L)[i * 3 + i] = std::sqrt(s);
                }
                else
                {
                    (*L)[i * 3 + j] = s / (*L)[j * 3 + j];
                }
            }
        }
        return true;
    }

    bool controlWhitening(int kind, double sigma, const double* m, Mat3* A)
    {
        *A = Mat3{};
        if (kind == PB_CONTROL_ISOTROPIC_SIGMA)
        {
            if (!std::isfinite(sigma) || sigma <= 0.0) return false;
            (*A)[0] = (*A)[4] = (*A)[8] = 1.0 / sigma;
            return true;
        }
        if (!m) return false;
        Mat3 M;
        std::copy(m, m + 9, M.begin());
        if (kind == PB_CONTROL_COVARIANCE)
        {
            Mat3 L;
            if (!cholesky3(M, &L)) return false;
            // A = L^-1 (lower triangular)
            (*A)[0] = 1.0 / L[0];
            (*A)[4] = 1.0 / L[4];
            (*A)[8] = 1.0 / L[8];
            (*A)[3] = -L[3] * (*A)[0] / L[4];
            (*A)[7] = -L[7] * (*A)[4] / L[8];
            (*A)[6] = -(L[6] * (*A)[0] + L[7] * (*A)[3]) / L[8];
            return true;
        }
        if (kind == PB_CONTROL_SQRT_INFORMATION)
        {
            Mat3 I{}, L;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) I[i * 3 + j] += M[k * 3 + i] * M[k * 3 + j];
            if (!cholesky3(I, &L)) return false;
            *A = M;
            return true;
        }
        return false;
    }

    // --- observation sanitising (BundleAdjustValidation.cpp:304-326) -------------------------
    double sanitizedWeight(const Observation& o)
    {
        return std::isfinite(o.weight) && o.weight > 0.0 ? o.weight : 0.0;
    }
    double sanitizedScale(const Observation& o)
    {
        return std::isfinite(o.scale) && o.scale > 0.0 ? o.scale : 1.0;
    }
    bool observationUsable(const Observation& o, std::size_t cameras)
    {
        return o.cam >= 0 && static_cast<std::size_t>(o.cam) < cameras && std::isfinite(o.u) &&
               std::isfinite(o.v) && sanitizedWeight(o) > 0.0;
    }

    // --- supported parameter mask (CameraState.cpp:187-204), perspective model only ---------
    SlotMask supportedSlots(int conv)
    {
        SlotMask m{};
        for (int i = 0; i <= 10; ++i) m[i] = true;
        m[11] = m[12] = conv == PB_BROWN_METASHAPE;
        return m;
    }

    // --- frame camera projection --------------------------------------------------------------
    // Pinhole looking along +Z with Brown distortion on the normalised plane.
    //   OpenCV:    xd = x*rad + 2 p1 x y + p2 (r2 + 2x^2);  yd = y*rad + p1 (r2 + 2y^2) + 2 p2 x y
    //   Metashape: xd = x*rad + (p1 (r2 + 2x^2) + 2 p2 x y)(1 + p3 r2 + p4 r2^2)
    //              yd = y*rad + (p2 (r2 + 2y^2) + 2 p1 x y)(1 + p3 r2 + p4 r2^2)
    //   rad = 1 + k1 r2 + k2 r2^2 + k3 r2^3 + k4 r2^4
    //   u = fx xd + b2 yd + cx,   v = fy yd + cy,   fy = fx exp(slot1)
    // Forward depth must exceed 1e-9 (BundleAdjustProjection.h:74-78).
    struct ProjectionOut
    {
        double uv[2];
        double depth;
        double Jcam[6];    // d(u,v)/d(camera-frame point), row-major 2x3
        double Jslot[26];  // d(u,v)/d(slot), row 0 = u (13), row 1 = v (13)
    };

    bool project(const Camera& cam, const Vec3& X, bool with_jacobian, ProjectionOut* out)
    {
        const double d0 = X[0] - cam.C[0], d1 = X[1] - cam.C[1], d2 = X[2] - cam.C[2];
        const double xc0 = cam.R[0] * d0 + cam.R[3] * d1 + cam.R[6] * d2;
        const double xc1 = cam.R[1] * d0 + cam.R[4] * d1 + cam.R[7] * d2;
        const double xc2 = cam.R[2] * d0 + cam.R[5] * d1 + cam.R[8] * d2;
        if (!(xc2 > 1e-9) || !std::isfinite(xc0) || !std::isfinite(xc1) || !std::isfinite(xc2))
        {
            return false;
        }
        const double iz = 1.0 / xc2;
        const double x = xc0 * iz, y = xc1 * iz;
        const Slots& s = cam.p;
        const double fx = s[0];
        const double fy = fx * std::exp(s[1]);
        const double k1 = s[4], k2 = s[5], k3 = s[6], tp1 = s[7], tp2 = s[8], b2 = s[9], k4 = s[10];
        const double tp3 = s[11], tp4 = s[12];
        const double r2 = x * x + y * y;
        const double rad = 1.0 + r2 * (k1 + r2 * (k2 + r2 * (k3 + r2 * k4)));
        const bool meta = cam.conv == PB_BROWN_METASHAPE;
        double tx0, ty0, T;
        if (meta)
        {
            tx0 = tp1 * (r2 + 2.0 * x * x) + 2.0 * tp2 * x * y;
            ty0 = tp2 * (r2 + 2.0 * y * y) + 2.0 * tp1 * x * y;
            T = 1.0 + tp3 * r2 + tp4 * r2 * r2;
        }
        else
        {
            tx0 = 2.0 * tp1 * x * y + tp2 * (r2 + 2.0 * x * x);
            ty0 = tp1 * (r2 + 2.0 * y * y) + 2.0 * tp2 * x * y;
            T = 1.0;
        }
        const double xd = x * rad + tx0 * T;
        const double yd = y * rad + ty0 * T;
        out->uv[0] = fx * xd + b2 * yd + s[2];
        out->uv[1] = fy * yd + s[3];
        out->depth = xc2;
        if (!std::isfinite(out->uv[0]) || !std::isfinite(out->uv[1]))
        {
            return false;
        }
        if (!with_jacobian)
        {
            return true;
        }
        // d(rad)/d(r2)
        const double drad = k1 + r2 * (2.0 * k2 + r2 * (3.0 * k3 + r2 * 4.0 * k4));
        double dtx_dx, dtx_dy, dty_dx, dty_dy;
        if (meta)
        {
            const double dT = tp3 + 2.0 * tp4 * r2;  // dT/dr2
            const double a_dx = 6.0 * tp1 * x + 2.0 * tp2 * y;
            const double a_dy = 2.0 * tp1 * y + 2.0 * tp2 * x;
            const double b_dx = 2.0 * tp2 * x + 2.0 * tp1 * y;
            const double b_dy = 6.0 * tp2 * y + 2.0 * tp1 * x;
            dtx_dx = a_dx * T + tx0 * dT * 2.0 * x;
            dtx_dy = a_dy * T + tx0 * dT * 2.0 * y;
            dty_dx = b_dx * T + ty0 * dT * 2.0 * x;
            dty_dy = b_dy * T + ty0 * dT * 2.0 * y;
        }
        else
        {
            dtx_dx = 2.0 * tp1 * y + 6.0 * tp2 * x;
            dtx_dy = 2.0 * tp1 * x + 2.0 * tp2 * y;
            dty_dx = 2.0 * tp1 * x + 2.0 * tp2 * y;
            dty_dy = 6.0 * tp1 * y + 2.0 * tp2 * x;
        }
        const double dxd_dx = rad + x * drad * 2.0 * x + dtx_dx;
        const double dxd_dy = x * drad * 2.0 * y + dtx_dy;
        const double dyd_dx = y * drad * 2.0 * x + dty_dx;
        const double dyd_dy = rad + y * drad * 2.0 * y + dty_dy;
        const double du_dx = fx * dxd_dx + b2 * dyd_dx;
        const double du_dy = fx * dxd_dy + b2 * dyd_dy;
        const double dv_dx = fy * dyd_dx;
        const double dv_dy = fy * dyd_dy;
        // (x, y) = (xc0/xc2, xc1/xc2)
        out->Jcam[0] = du_dx * iz;
        out->Jcam[1] = du_dy * iz;
        out->Jcam[2] = -(du_dx * x + du_dy * y) * iz;
        out->Jcam[3] = dv_dx * iz;
        out->Jcam[4] = dv_dy * iz;
        out->Jcam[5] = -(dv_dx * x + dv_dy * y) * iz;

        double* Ju = out->Jslot;
        double* Jv = out->Jslot + kP;
        std::fill(Ju, Ju + 2 * kP, 0.0);
        const double aspect = fy / fx;
        // focal at fixed aspect and log-aspect (chain rule of CameraState.cpp:468-473)
        Ju[0] = xd;
        Jv[0] = aspect * yd;
        Ju[1] = 0.0;
        Jv[1] = fy * yd;
        Ju[2] = 1.0;
        Jv[3] = 1.0;
        auto dist = [&](int slot, double dxd, double dyd)
        {
            Ju[slot] = fx * dxd + b2 * dyd;
            Jv[slot] = fy * dyd;
        };
        const double r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4;
        dist(4, x * r2, y * r2);
        dist(5, x * r4, y * r4);
        dist(6, x * r6, y * r6);
        dist(10, x * r8, y * r8);
        if (meta)
        {
            dist(7, (r2 + 2.0 * x * x) * T, 2.0 * x * y * T);
            dist(8, 2.0 * x * y * T, (r2 + 2.0 * y * y) * T);
            dist(11, tx0 * r2, ty0 * r2);
            dist(12, tx0 * r4, ty0 * r4);
        }
        else
        {
            dist(7, 2.0 * x * y, r2 + 2.0 * y * y);
            dist(8, r2 + 2.0 * x * x, 2.0 * x * y);
        }
        Ju[9] = yd;
        Jv[9] = 0.0;
        return true;
    }

    // --- robust loss (BundleAdjustPlaMatrixProjection.cpp:75-109) -----------------------------
    bool robustLoss(double s, int loss, double scale, double* cost, double* weight)
    {
        if (!std::isfinite(s) || s < 0.0 || !std::isfinite(scale) || scale <= 0.0)
        {
            return false;
        }
        const double d2 = scale * scale;
        switch (loss)
        {
        case PB_LOSS_LEAST_SQUARES:
            *cost = 0.5 * s;
            *weight = 1.0;
            return true;
        case PB_LOSS_HUBER:
            if (s <= d2)
            {
                *cost = 0.5 * s;
                *weight = 1.0;
                return true;
            }
            *cost = scale * std::sqrt(s) - 0.5 * d2;
            *weight = scale / std::sqrt(s);
            return true;
        case PB_LOSS_CAUCHY:
            *cost = 0.5 * d2 * std::log1p(s / d2);
            *weight = 1.0 / (1.0 + s / d2);
            return true;
        default:
            return false;
        }
    }

    // --- point scale of the reference parameterisation (Projection.cpp:15-24) -----------------
    double pointScale(const Vec3& X)
    {
        const double len = std::sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        const double s = len / 1000.0;
        if (!std::isfinite(s) || s == 0.0)
        {
            return 1.0;
        }
        return s / std::atan(s);
    }

    Mat3 mul3(const Mat3& a, const Mat3& b)
    {
        Mat3 c{};
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col)
                for (int k = 0; k < 3; ++k) c[r * 3 + col] += a[r * 3 + k] * b[k * 3 + col];
        return c;
    }

    // Reference pose retraction: C2W' = D E^T D C2W, E = Rx Ry Rz (Assembly.cpp:130-179).
    void retractReference(Mat3* R, Vec3* C, const double* delta)
    {
        const double sx = std::sin(delta[0]), cx = std::cos(delta[0]);
        const double sy = std::sin(delta[1]), cy = std::cos(delta[1]);
        const double sz = std::sin(delta[2]), cz = std::cos(delta[2]);
        const Mat3 rx{{1, 0, 0, 0, cx, -sx, 0, sx, cx}};
        const Mat3 ry{{cy, 0, sy, 0, 1, 0, -sy, 0, cy}};
        const Mat3 rz{{cz, -sz, 0, sz, cz, 0, 0, 0, 1}};
        const Mat3 e = mul3(mul3(rx, ry), rz);
        const double sg[3] = {1.0, -1.0, -1.0};
        Mat3 left{};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) left[r * 3 + c] = sg[r] * e[c * 3 + r] * sg[c];
        *R = mul3(left, *R);
        (*C)[0] += delta[3];
        (*C)[1] += delta[4];
        (*C)[2] += delta[5];
    }

    // Left exponential-map retraction R' = Exp(w) R, C' = C + dC (BundleAdjustProjection.h:150-231).
    void retractExp(Mat3* R, Vec3* C, const double* delta)
    {
        const double wx = delta[0], wy = delta[1], wz = delta[2];
        const double t2 = wx * wx + wy * wy + wz * wz;
        Mat3 d{};
        if (t2 < 1e-20)
        {
            d = {1.0, -wz, wy, wz, 1.0, -wx, -wy, wx, 1.0};
        }
        else
        {
            const double t = std::sqrt(t2);
            const double sc = std::sin(t) / t, cc = (1.0 - std::cos(t)) / t2;
            d = {1.0 - cc * (wy * wy + wz * wz), cc * wx * wy - sc * wz, cc * wx * wz + sc * wy,
                 cc * wx * wy + sc * wz, 1.0 - cc * (wx * wx + wz * wz), cc * wy * wz - sc * wx,
                 cc * wx * wz - sc * wy, cc * wy * wz + sc * wx, 1.0 - cc * (wx * wx + wy * wy)};
        }
        *R = mul3(d, *R);
        (*C)[0] += delta[3];
        (*C)[1] += delta[4];
        (*C)[2] += delta[5];
    }

    // --- per-observation linearisation (Projection.cpp:176-285) -------------------------------
    struct Linearization
    {
        double r[2];
        double Jp[6];     // 2x3 world point Jacobian (scaled by the point scale when reference param)
        double Jpose[12]; // 2x6 [rotation, centre]
        double Jint[26];  // 2x13 slot Jacobian (active and supported slots only)
        double w;         // normal weight = whitened weight * robust weight
        double cost;
    };

    bool linearize(const Camera& cam, const Vec3& X, const Observation& o, int loss, double lscale,
                   bool reference_param, const SlotMask& active, Linearization* L)
    {
        ProjectionOut pr;
        if (!project(cam, X, true, &pr))
        {
            return false;
        }
        L->r[0] = pr.uv[0] - o.u;
        L->r[1] = pr.uv[1] - o.v;
        const Mat3& R = cam.R;
        for (int a = 0; a < 2; ++a)
        {
            for (int w = 0; w < 3; ++w)
            {
                double dv = 0.0;
                for (int c = 0; c < 3; ++c) dv += pr.Jcam[a * 3 + c] * R[w * 3 + c];
                L->Jp[a * 3 + w] = dv;
                L->Jpose[a * 6 + 3 + w] = -dv;
            }
            const double* jp = L->Jp + a * 3;
            const double dx = X[0] - cam.C[0], dy = X[1] - cam.C[1], dz = X[2] - cam.C[2];
            if (!reference_param)
            {
                L->Jpose[a * 6 + 0] = jp[1] * dz - jp[2] * dy;
                L->Jpose[a * 6 + 1] = -jp[0] * dz + jp[2] * dx;
                L->Jpose[a * 6 + 2] = jp[0] * dy - jp[1] * dx;
            }
            else
            {
                // target frame D = diag(1,-1,-1): J_i = sum_t Jp_t * s_t * [D d]x(i, t)
                const double td[3] = {dx, -dy, -dz};
                const double tc[3][3] = {{0.0, -td[2], td[1]}, {td[2], 0.0, -td[0]}, {-td[1], td[0], 0.0}};
                const double sg[3] = {1.0, -1.0, -1.0};
                for (int i = 0; i < 3; ++i)
                {
                    double v = 0.0;
                    for (int t = 0; t < 3; ++t) v += jp[t] * sg[t] * tc[i][t];
                    L->Jpose[a * 6 + i] = v;
                }
            }
        }
        if (reference_param)
        {
            const double ps = pointScale(X);
            for (double& v : L->Jp) v *= ps;
        }
        double w = sanitizedWeight(o);
        const double ms = sanitizedScale(o);
        w /= ms * ms;
        const double s = w * (L->r[0] * L->r[0] + L->r[1] * L->r[1]);
        double cost = 0.0, rw = 0.0;
        if (!robustLoss(s, loss, lscale, &cost, &rw))
        {
            return false;
        }
        L->w = w * rw;
        L->cost = cost;
        const SlotMask sup = supportedSlots(cam.conv);
        for (int i = 0; i < kP; ++i)
        {
            const bool on = active[i] && sup[i];
            L->Jint[i] = on ? pr.Jslot[i] : 0.0;
            L->Jint[kP + i] = on ? pr.Jslot[kP + i] : 0.0;
        }
        for (double v : L->Jint)
            if (!std::isfinite(v)) return false;
        return std::isfinite(L->w) && std::isfinite(L->cost);
    }

    bool observationCost(const Camera& cam, const Vec3& X, const Observation& o, int loss, double lscale,
                         double* cost)
    {
        ProjectionOut pr;
        if (!project(cam, X, false, &pr))
        {
            return false;
        }
        const double rx = pr.uv[0] - o.u, ry = pr.uv[1] - o.v;
        double w = sanitizedWeight(o);
        const double ms = sanitizedScale(o);
        w /= ms * ms;
        double c = 0.0, rw = 0.0;
        if (!robustLoss(w * (rx * rx + ry * ry), loss, lscale, &c, &rw))
        {
            return false;
        }
        *cost = c;
        return std::isfinite(c);
    }

    // --- survey constraints (BundleAdjustPlaMatrixConstraints.cpp:17-29, 66-116) ----------------
    // Residual r = J (X - target) with J = sqrt(w) A (control point, 3 rows) or sqrt(w) n^T (laser plane, 1
    // row), robustified by Huber on |r|^2: delta_cp for control points, delta_l * sqrt(w) for laser planes.
    struct ConstraintLin
    {
        int rows = 0;
        double r[3] = {0.0, 0.0, 0.0};
        double J[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double w = 0.0;
        double cost = 0.0;
    };

    bool finishConstraint(ConstraintLin* c, double delta)
    {
        double s = 0.0;
        for (int i = 0; i < c->rows; ++i) s += c->r[i] * c->r[i];
        double cost = 0.0, w = 0.0;
        if (!robustLoss(s, PB_LOSS_HUBER, delta, &cost, &w)) return false;
        c->w = w;
        c->cost = cost;
        return std::isfinite(w) && std::isfinite(cost);
    }

    bool linearizeControl(const ControlPoint& k, const Vec3& X, const pb_solve_options& o, ConstraintLin* c)
    {
        const double w = o.control_point_weight * k.weight;
        if (!(w > 0.0)) return false;
        const double sw = std::sqrt(w);
        *c = ConstraintLin{};
        c->rows = 3;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                c->J[i * 3 + j] = sw * k.A[i * 3 + j];
                c->r[i] += c->J[i * 3 + j] * (X[j] - k.P[j]);
            }
        return finishConstraint(c, o.control_point_huber_delta);
    }

    bool linearizePlane(const LaserPlane& k, const Vec3& X, const pb_solve_options& o, ConstraintLin* c)
    {
        const double w = o.laser_plane_weight * k.weight;
        if (!(w > 0.0)) return false;
        const double sw = std::sqrt(w);
        *c = ConstraintLin{};
        c->rows = 1;
        for (int j = 0; j < 3; ++j)
        {
            c->J[j] = sw * k.n[j];
            c->r[0] += c->J[j] * (X[j] - k.p[j]);
        }
        return finishConstraint(c, o.laser_huber_delta * sw);
    }

    // --- problem, options, active problem -----------------------------------------------------
    struct Options
    {
        pb_solve_options o;
        std::vector<int> fixedCameras;
        std::vector<int> fixedTracks;  // sorted
        int gaugeAnchor = -1;
        int gaugeScale = -1;
        double gaugeBaseline = 0.0;
        int gaugePolicy = PB_GAUGE_AUTO_ANCHOR;
        bool controlPoints = false;  // enableControlPointConstraints (Conversion.cpp:106)
        bool laserPlanes = false;    // enableLaserPlaneConstraints (Conversion.cpp:103)
        bool rig = false;            // !options.rig.empty() (Conversion.cpp:144)
        // useReferencePointParameterization (Assembly.cpp:24-29)
        bool referenceParam() const { return o.use_online_schur != 0 && !rig && !controlPoints && !laserPlanes; }
        bool enabled(int slot) const { return (o.shared_intrinsic_mask >> slot) & 1u; }
    };

    struct Active
    {
        std::vector<char> activeTrack;
        std::vector<int> cameraBlock, intrinsicBlockByCamera, calibrationGroupByCamera, trackBlock;
        std::vector<int> rigCaptureBlockByCamera, rigSensorBlockByCamera;
        std::vector<int> rigCaptureBlock, rigSensorBlock;  // per capture / sensor, -1 when not refined
        int primaryBlockCount = 0, cameraBlockCount = 0, intrinsicBlockCount = 0, trackBlockCount = 0;
        int rigCaptureBlockCount = 0, rigSensorBlockCount = 0;
        int activeTrackCount = 0, observationCount = 0, rejectedInitialTracks = 0;
    };

    struct Group
    {
        Slots params{}, prior{}, invSigma{}, lower{}, upper{};
        SlotMask enabled{};
        double focalRef = 1.0, aspectRef = 1.0;
    };

    // Rig poses of the state (OptimizationState::rig, BundleAdjustPlaMatrixAssembly.h:27): captures carry the
    // rig-to-world rotation and the rig centre, sensors the camera-to-rig rotation and the camera centre in the
    // rig frame.
    struct RigPoses
    {
        std::vector<Mat3> capR, senR;
        std::vector<Vec3> capC, senC;
    };

    struct State
    {
        std::vector<Camera> cams;
        std::vector<Vec3> points;
        std::vector<Group> groups;
        RigPoses rig;
    };

    // finiteMedian: unpinned choice - odd count: middle value; even count: mean of the two middle values.
    double medianOr(std::vector<double> v, double fallback)
    {
        v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return !std::isfinite(x); }), v.end());
        if (v.empty()) return fallback;
        std::sort(v.begin(), v.end());
        const std::size_t n = v.size();
        return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
    }

    // Rig topology of the problem: poses, fixed flags and per camera its binding (capture, sensor), -1 when
    // unbound.
    struct Rig
    {
        RigPoses poses;
        std::vector<char> capFixed, senFixed;
        std::vector<int> camCapture, camSensor;
        bool empty() const { return poses.capR.empty() && poses.senR.empty(); }
        bool bound(int c) const
        {
            return !empty() && c >= 0 && c < static_cast<int>(camCapture.size()) && camCapture[c] >= 0 &&
                   camSensor[c] >= 0;
        }
    };

    // composeRigCameras (placamera, called at solver.cpp:282 and Assembly.cpp:1203): a bound camera's
    // camera-to-world rotation is R_rig R_sensor and its centre C_rig + R_rig c_sensor. The composition is
    // PlaCamera's and not in the tree; this form is the one the rig Jacobians of cameraPosePrimaryTerms
    // (Assembly.cpp:224-281) differentiate: the lever arm R_rig c_sensor and the rig-frame sensor axes R_rig.
    void composeRig(const Rig& rig, const RigPoses& poses, std::vector<Camera>* cams)
    {
        for (std::size_t c = 0; c < cams->size(); ++c)
        {
            if (!rig.bound(static_cast<int>(c))) continue;
            const int k = rig.camCapture[c], s = rig.camSensor[c];
            Camera& cam = (*cams)[c];
            cam.R = mul3(poses.capR[k], poses.senR[s]);
            for (int i = 0; i < 3; ++i)
            {
                double v = poses.capC[k][i];
                for (int j = 0; j < 3; ++j) v += poses.capR[k][i * 3 + j] * poses.senC[s][j];
                cam.C[i] = v;
            }
        }
    }

    struct Problem
    {
        Rig rig;
        std::vector<Camera> cams;
        std::vector<Slots> refIntr;
        std::vector<int> groupIds;
        bool hasGroupIds = false;
        std::vector<Track> tracks;
        std::vector<int> fixedCameras;
        std::vector<int> fixedTracks;
        int gaugePolicy = PB_GAUGE_AUTO_ANCHOR;
        int gaugeAnchor = -1, gaugeScale = -1;
        double gaugeBaseline = 0.0;
        std::string constraintError;  // validateProblem's verdict on the survey constraints (problem.cpp:241-259)
        std::string rigError;         // the rig could not be composed (runConcreteBackend, solver.cpp:282-287)
    };

    bool hasSharedIntrinsics(const Options& opt)
    {
        for (int i = 0; i < kP; ++i)
            if (opt.enabled(i)) return true;
        return false;
    }

    // hasUsableTrackObservations (BundleAdjustValidation.cpp:94-106): two distinct cameras with a finite
    // pixel and a positive weight (the camera index is only checked for being non-negative).
    bool usableForGauge(const Track& tr)
    {
        std::set<int> cams;
        for (const Observation& ob : tr.obs)
            if (ob.cam >= 0 && std::isfinite(ob.u) && std::isfinite(ob.v) && std::isfinite(ob.weight) &&
                ob.weight > 0.0)
                cams.insert(ob.cam);
        return cams.size() >= 2;
    }

    // hasCompleteControlPointGauge (BundleAdjustValidation.cpp:181-240): three non-collinear targets.
    bool completeControlGauge(const Problem& pb)
    {
        std::vector<Vec3> pts;
        for (const Track& tr : pb.tracks)
        {
            if (tr.controls.empty() || !usableForGauge(tr)) continue;
            for (const ControlPoint& k : tr.controls) pts.push_back(k.P);
        }
        for (std::size_t a = 0; a < pts.size(); ++a)
            for (std::size_t b = a + 1; b < pts.size(); ++b)
            {
                const Vec3 ab{pts[b][0] - pts[a][0], pts[b][1] - pts[a][1], pts[b][2] - pts[a][2]};
                const double ab2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
                if (ab2 <= 1e-16) continue;
                for (std::size_t c = b + 1; c < pts.size(); ++c)
                {
                    const Vec3 ac{pts[c][0] - pts[a][0], pts[c][1] - pts[a][1], pts[c][2] - pts[a][2]};
                    const double ac2 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
                    const Vec3 x{ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                                 ab[0] * ac[1] - ab[1] * ac[0]};
                    const double x2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                    if (ac2 > 1e-16 && x2 > 1e-12 * ab2 * ac2) return true;
                }
            }
        return false;
    }

    // hasControlPointBaselineToFixedCamera (BundleAdjustValidation.cpp:147-179).
    bool controlBaselineToFixed(const Problem& pb, const std::set<int>& fixed)
    {
        for (const Track& tr : pb.tracks)
        {
            if (tr.controls.empty() || !usableForGauge(tr)) continue;
            for (const ControlPoint& k : tr.controls)
                for (int f : fixed)
                {
                    const Vec3& c = pb.cams[f].C;
                    if (!std::isfinite(c[0]) || !std::isfinite(c[1]) || !std::isfinite(c[2])) continue;
                    const double d = (k.P[0] - c[0]) * (k.P[0] - c[0]) + (k.P[1] - c[1]) * (k.P[1] - c[1]) +
                                     (k.P[2] - c[2]) * (k.P[2] - c[2]);
                    if (d > 1e-16) return true;
                }
        }
        return false;
    }

    // validateAndNormalizeBundleAdjustOptions gauge part (BundleAdjustValidation.cpp:767-896) for
    // problems without pose priors or scale bars: control points (three non-collinear targets) give
    // absolute pose and scale; a target away from a fixed camera's centre gives absolute scale.
    std::string normalizeGauge(const Problem& pb, Options* opt)
    {
        const int n = static_cast<int>(pb.cams.size());
        std::set<int> fixed;
        for (int idx : pb.fixedCameras)
        {
            if (idx < 0 || idx >= n) return "fixed camera index out of range";
            if (!fixed.insert(idx).second) return "duplicate fixed camera index";
        }
        const bool requested = pb.gaugeAnchor >= 0 || pb.gaugeScale >= 0 || pb.gaugeBaseline > 0.0;
        if (requested && (pb.gaugeAnchor < 0 || pb.gaugeScale < 0 || pb.gaugeAnchor >= n || pb.gaugeScale >= n ||
                          pb.gaugeAnchor == pb.gaugeScale || !std::isfinite(pb.gaugeBaseline) ||
                          pb.gaugeBaseline <= 0.0))
            return "invalid reference gauge";
        std::set<int> fixedTracks;
        for (int idx : pb.fixedTracks)
        {
            if (idx < 0 || idx >= static_cast<int>(pb.tracks.size())) return "fixed track index out of range";
            if (!fixedTracks.insert(idx).second) return "duplicate fixed track index";
        }
        opt->fixedCameras = pb.fixedCameras;
        opt->fixedTracks.assign(fixedTracks.begin(), fixedTracks.end());
        opt->gaugeAnchor = pb.gaugeAnchor;
        opt->gaugeScale = pb.gaugeScale;
        opt->gaugeBaseline = pb.gaugeBaseline;
        opt->gaugePolicy = pb.gaugePolicy;
        if (!opt->o.refine_camera_pose || pb.gaugePolicy == PB_GAUGE_CALLER_MANAGED || n == 0) return {};
        const bool complete = opt->controlPoints && completeControlGauge(pb);
        const bool absolutePose = complete;
        bool absoluteScale = complete || (opt->controlPoints && controlBaselineToFixed(pb, fixed));
        if (pb.gaugePolicy == PB_GAUGE_REQUIRE_EXPLICIT)
        {
            if ((absolutePose && absoluteScale) || fixed.size() >= 2 || (fixed.size() == 1 && absoluteScale))
                return {};
            return "unsupported: incomplete gauge (fix two cameras, or one camera with absolute scale)";
        }
        if (absolutePose && absoluteScale) return {};
        if (fixed.empty())
        {
            opt->fixedCameras.push_back(0);
            fixed.insert(0);
            absoluteScale = absoluteScale || (opt->controlPoints && controlBaselineToFixed(pb, fixed));
        }
        if (!absoluteScale && fixed.size() < 2)
        {
            const int anchor = *fixed.begin();
            const Vec3 a = pb.cams[anchor].C;
            int best = -1;
            double bestD = 1e-20;
            for (int i = 0; i < n; ++i)
            {
                if (fixed.count(i)) continue;
                const Vec3 c = pb.cams[i].C;
                const double dd = (c[0] - a[0]) * (c[0] - a[0]) + (c[1] - a[1]) * (c[1] - a[1]) +
                                  (c[2] - a[2]) * (c[2] - a[2]);
                if (std::isfinite(dd) && dd > bestD)
                {
                    bestD = dd;
                    best = i;
                }
            }
            if (best < 0) return "unsupported: no second camera with a non-zero baseline";
            const Vec3 c = pb.cams[best].C;
            opt->gaugeAnchor = anchor;
            opt->gaugeScale = best;
            opt->gaugeBaseline = std::sqrt((c[0] - a[0]) * (c[0] - a[0]) + (c[1] - a[1]) * (c[1] - a[1]) +
                                           (c[2] - a[2]) * (c[2] - a[2]));
        }
        return {};
    }

    // prepareActiveProblem (BundleAdjustPlaMatrixProblem.cpp:29-278), frame cameras without rigs.
    Active prepareActive(const Problem& pb, const Options& opt)
    {
        Active a;
        const std::size_t nc = pb.cams.size();
        a.activeTrack.assign(pb.tracks.size(), 0);
        a.cameraBlock.assign(nc, -1);
        a.intrinsicBlockByCamera.assign(nc, -1);
        a.calibrationGroupByCamera.assign(nc, 0);
        a.trackBlock.assign(pb.tracks.size(), -1);
        std::vector<char> hasResidual(nc, 0);
        for (std::size_t t = 0; t < pb.tracks.size(); ++t)
        {
            const Track& tr = pb.tracks[t];
            if (!std::isfinite(tr.X[0]) || !std::isfinite(tr.X[1]) || !std::isfinite(tr.X[2])) continue;
            int first = -1;
            bool second = false, positive = true;
            int count = 0;
            double sq = 0.0;
            for (const Observation& o : tr.obs)
            {
                if (!observationUsable(o, nc)) continue;
                first = first < 0 ? o.cam : first;
                second = second || o.cam != first;
                ProjectionOut pr;
                if (!project(pb.cams[o.cam], tr.X, false, &pr))
                {
                    positive = false;
                    break;
                }
                const double du = pr.uv[0] - o.u, dv = pr.uv[1] - o.v;
                sq += sanitizedWeight(o) * (du * du + dv * dv);
                ++count;
            }
            const double rms = count > 0 ? std::sqrt(sq / (2.0 * count)) : kInf;
            if (count < 2 || !second || !positive) continue;
            if (opt.o.max_initial_track_rms > 0.0 && rms > opt.o.max_initial_track_rms)
            {
                ++a.rejectedInitialTracks;
                continue;
            }
            a.activeTrack[t] = 1;
            ++a.activeTrackCount;
            a.observationCount += count;
            for (const Observation& o : tr.obs)
                if (observationUsable(o, nc)) hasResidual[o.cam] = 1;
        }
        auto isFixed = [&](int c)
        { return std::find(opt.fixedCameras.begin(), opt.fixedCameras.end(), c) != opt.fixedCameras.end(); };
        const Rig& rig = pb.rig;
        a.rigCaptureBlockByCamera.assign(nc, -1);
        a.rigSensorBlockByCamera.assign(nc, -1);
        a.rigCaptureBlock.assign(rig.poses.capR.size(), -1);
        a.rigSensorBlock.assign(rig.poses.senR.size(), -1);
        if (opt.o.refine_camera_pose && rig.empty())
        {
            for (std::size_t c = 0; c < nc; ++c)
                if (hasResidual[c] && !isFixed(static_cast<int>(c))) a.cameraBlock[c] = a.cameraBlockCount++;
        }
        else if (opt.o.refine_camera_pose)
        {
            // BundleAdjustPlaMatrixProblem.cpp:157-199: a capture is refined when one of its cameras has a residual
            // and it is not fixed, none of its cameras is fixed and none is the gauge scale camera; a sensor when
            // one of its cameras has a residual and its extrinsic is not fixed; sensor blocks follow the captures
            for (std::size_t k = 0; k < rig.poses.capR.size(); ++k)
            {
                bool residual = false, fixed = rig.capFixed[k] != 0;
                for (std::size_t c = 0; c < nc; ++c)
                    if (rig.camCapture[c] == static_cast<int>(k))
                    {
                        residual = residual || hasResidual[c];
                        fixed = fixed || isFixed(static_cast<int>(c)) || opt.gaugeScale == static_cast<int>(c);
                    }
                if (residual && !fixed) a.rigCaptureBlock[k] = a.rigCaptureBlockCount++;
            }
            for (std::size_t k = 0; k < rig.poses.senR.size(); ++k)
            {
                bool residual = false;
                for (std::size_t c = 0; c < nc; ++c)
                    residual = residual || (rig.camSensor[c] == static_cast<int>(k) && hasResidual[c]);
                if (residual && !rig.senFixed[k])
                    a.rigSensorBlock[k] = a.rigCaptureBlockCount + a.rigSensorBlockCount++;
            }
            for (std::size_t c = 0; c < nc; ++c)
            {
                if (rig.camCapture[c] >= 0) a.rigCaptureBlockByCamera[c] = a.rigCaptureBlock[rig.camCapture[c]];
                if (rig.camSensor[c] >= 0) a.rigSensorBlockByCamera[c] = a.rigSensorBlock[rig.camSensor[c]];
            }
            a.cameraBlockCount = a.rigCaptureBlockCount + a.rigSensorBlockCount;
        }
        a.primaryBlockCount = a.cameraBlockCount;
        if (hasSharedIntrinsics(opt))
        {
            if (pb.hasGroupIds)
                for (std::size_t c = 0; c < nc; ++c) a.calibrationGroupByCamera[c] = pb.groupIds[c];
            std::vector<int> ids;
            for (std::size_t c = 0; c < nc; ++c)
            {
                const int id = a.calibrationGroupByCamera[c];
                auto it = std::find(ids.begin(), ids.end(), id);
                int gi;
                if (it == ids.end())
                {
                    gi = static_cast<int>(ids.size());
                    ids.push_back(id);
                }
                else
                {
                    gi = static_cast<int>(it - ids.begin());
                }
                a.calibrationGroupByCamera[c] = gi;
            }
            a.intrinsicBlockCount = static_cast<int>(ids.size());
            for (std::size_t c = 0; c < nc; ++c)
                a.intrinsicBlockByCamera[c] = a.primaryBlockCount + a.calibrationGroupByCamera[c];
            a.primaryBlockCount += a.intrinsicBlockCount;
        }
        for (std::size_t t = 0; t < pb.tracks.size(); ++t)
        {
            if (!a.activeTrack[t]) continue;
            if (std::binary_search(opt.fixedTracks.begin(), opt.fixedTracks.end(), static_cast<int>(t))) continue;
            a.trackBlock[t] = a.trackBlockCount++;
        }
        return a;
    }

    // initializeIntrinsicGroups (BundleAdjustPlaMatrixModel.cpp:84-231) without transition priors.
    std::vector<Group> initGroups(const Problem& pb, const Options& opt, const Active& a)
    {
        std::vector<Group> groups(static_cast<std::size_t>(a.intrinsicBlockCount));
        if (groups.empty()) return groups;
        struct Samples
        {
            std::array<std::vector<double>, kP> init, ref;
            std::vector<double> widths, heights;
            SlotMask sup;
        };
        std::vector<Samples> smp(groups.size());
        for (auto& s : smp) s.sup.fill(true);
        for (std::size_t c = 0; c < pb.cams.size(); ++c)
        {
            Samples& s = smp[a.calibrationGroupByCamera[c]];
            const Slots& p = pb.cams[c].p;
            const Slots& r = pb.refIntr[c];
            s.init[0].push_back(p[0]);
            s.ref[0].push_back(r[0]);
            s.init[1].push_back(std::exp(p[1]));
            s.ref[1].push_back(std::exp(r[1]));
            s.init[2].push_back(p[2] - r[2]);
            s.init[3].push_back(p[3] - r[3]);
            s.ref[2].push_back(0.0);
            s.ref[3].push_back(0.0);
            for (int i = 4; i < kP; ++i)
            {
                s.init[i].push_back(p[i]);
                s.ref[i].push_back(r[i]);
            }
            const SlotMask sup = supportedSlots(pb.cams[c].conv);
            for (int i = 0; i < kP; ++i) s.sup[i] = s.sup[i] && sup[i];
            // appendSamples (Model.cpp:61-66): the image size, else twice the reference principal point
            const bool sized = pb.cams[c].width > 0 || pb.cams[c].height > 0;
            s.widths.push_back(sized ? pb.cams[c].width : 2.0 * r[2]);
            s.heights.push_back(sized ? pb.cams[c].height : 2.0 * r[3]);
        }
        const pb_solve_options& o = opt.o;
        const double low = std::max(1.0, o.low_order_distortion_scale);
        for (std::size_t gi = 0; gi < groups.size(); ++gi)
        {
            Group& g = groups[gi];
            g.lower.fill(-kInf);
            g.upper.fill(kInf);
            for (int i = 0; i < kP; ++i) g.enabled[i] = opt.enabled(i) && smp[gi].sup[i];
            g.focalRef = medianOr(smp[gi].ref[0], 1.0);
            g.aspectRef = medianOr(smp[gi].ref[1], 1.0);
            g.params[0] = std::max(1e-12, medianOr(smp[gi].init[0], g.focalRef));
            g.params[1] = std::log(std::max(1e-12, medianOr(smp[gi].init[1], g.aspectRef)));
            g.prior[0] = g.focalRef;
            g.prior[1] = std::log(g.aspectRef);
            for (int i = 2; i < kP; ++i)
            {
                g.prior[i] = medianOr(smp[gi].ref[i], 0.0);
                g.params[i] = medianOr(smp[gi].init[i], g.prior[i]);
            }
            g.lower[0] = g.focalRef * o.min_focal_scale;
            g.upper[0] = g.focalRef * o.max_focal_scale;
            g.lower[1] = std::log(g.aspectRef * o.min_aspect_scale);
            g.upper[1] = std::log(g.aspectRef * o.max_aspect_scale);
            const double pl = g.focalRef * o.max_principal_offset_fraction;
            g.lower[2] = g.lower[3] = -pl;
            g.upper[2] = g.upper[3] = pl;
            const double lim[9] = {o.max_abs[0] * low, o.max_abs[1], o.max_abs[2], o.max_abs[3] * low,
                                   o.max_abs[4] * low, g.focalRef * o.max_abs[5], o.max_abs[6], o.max_abs[7],
                                   o.max_abs[8]};
            for (int i = 0; i < 9; ++i)
            {
                g.lower[4 + i] = -lim[i];
                g.upper[4 + i] = lim[i];
            }
            const double psig = std::max(1e-6, g.focalRef * o.principal_prior_sigma_fraction);
            const double sig[kP] = {o.focal_prior_sigma * g.focalRef, o.aspect_prior_sigma, psig, psig,
                                    o.prior_sigma[0] * low, o.prior_sigma[1], o.prior_sigma[2],
                                    o.prior_sigma[3] * low, o.prior_sigma[4] * low,
                                    g.focalRef * o.prior_sigma[5], o.prior_sigma[6], o.prior_sigma[7],
                                    o.prior_sigma[8]};
            for (int i = 0; i < kP; ++i)
            {
                g.params[i] = std::clamp(g.params[i], g.lower[i], g.upper[i]);
                g.invSigma[i] = g.enabled[i] ? 1.0 / std::max(1e-9, sig[i]) : 0.0;
            }
            // useReferenceCalibrationTransitionPrior (Model.cpp:173-225). The transition weight replaces the
            // inverse sigma of every slot but the aspect; its reference value is the prior (prior[0] is the focal
            // reference), so the prior terms keep their form with this weight.
            if (o.use_transition_prior)
            {
                bool changed = false;
                for (int i = 0; i < kP; ++i)
                    if (i != 1) changed = changed || g.enabled[i] != (((o.previous_intrinsic_mask >> i) & 1u) != 0);
                const double w = medianOr(smp[gi].widths, 0.0), h = medianOr(smp[gi].heights, 0.0);
                const double nx = 0.5 * w / g.focalRef, ny = 0.5 * h / (g.focalRef * g.aspectRef);
                const double r = std::sqrt(nx * nx + ny * ny);
                Slots tw{};
                if (changed && std::isfinite(r) && r > 0.0)
                {
                    const double f = g.focalRef;
                    const Slots base{1.0 / r, 0.0, 1.0, 1.0, 1.0 / (std::pow(r, 3) * f), 1.0 / (std::pow(r, 5) * f),
                                     1.0 / (std::pow(r, 7) * f), 1.0 / (r * r * f), 1.0 / (r * r * f), 1.0 / r,
                                     1.0 / (std::pow(r, 9) * f), 1.0, 1.0};
                    for (int i = 0; i < kP; ++i)
                        if (i != 1 && g.enabled[i])
                            tw[i] = 1.0 / (base[i] * (((o.previous_intrinsic_mask >> i) & 1u) ? 1e6 : 0.01));
                }
                for (int i = 0; i < kP; ++i)
                    if (i != 1) g.invSigma[i] = tw[i];
            }
        }
        return groups;
    }

    // makeCameraWithSharedIntrinsics (BundleAdjustPlaMatrixProjection.cpp:26-63).
    Camera effectiveCamera(const Camera& cam, const Slots& ref, const Group& g)
    {
        Camera e = cam;
        const double fx = g.enabled[0] ? g.params[0] : cam.p[0];
        const double srcAspect = cam.p[0] > 1e-12 ? std::exp(cam.p[1]) : 1.0;
        const double aspect = g.enabled[1] ? std::exp(g.params[1]) : srcAspect;
        e.p[0] = fx;
        e.p[1] = std::log(aspect);
        e.p[2] = g.enabled[2] ? ref[2] + g.params[2] : cam.p[2];
        e.p[3] = g.enabled[3] ? ref[3] + g.params[3] : cam.p[3];
        const SlotMask sup = supportedSlots(cam.conv);
        for (int i = 4; i < kP; ++i)
            if (g.enabled[i] && sup[i]) e.p[i] = g.params[i];
        return e;
    }

    std::vector<Camera> effectiveCameras(const Problem& pb, const Active& a, const State& st)
    {
        if (st.groups.empty()) return st.cams;
        std::vector<Camera> out(st.cams.size());
        for (std::size_t c = 0; c < st.cams.size(); ++c)
            out[c] = effectiveCamera(st.cams[c], pb.refIntr[c], st.groups[a.calibrationGroupByCamera[c]]);
        return out;
    }

    // Reference gauge tangent basis (Assembly.cpp:59-128).
    Vec3 targetCoords(const Vec3& v) { return {v[0], -v[1], -v[2]}; }
    Vec3 cross(const Vec3& a, const Vec3& b)
    {
        return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    }
    double norm3(const Vec3& v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

    bool usesGaugeTangent(const Options& opt, const State& st, int cam)
    {
        const int n = static_cast<int>(st.cams.size());
        return opt.gaugeAnchor >= 0 && opt.gaugeScale >= 0 && opt.gaugeAnchor < n && opt.gaugeScale < n &&
               cam == opt.gaugeScale && std::isfinite(opt.gaugeBaseline) && opt.gaugeBaseline > 0.0;
    }

    std::array<Vec3, 2> gaugeBasis(const Options& opt, const State& st)
    {
        const Vec3 origin = targetCoords(st.cams[opt.gaugeAnchor].C);
        const Vec3 center = targetCoords(st.cams[opt.gaugeScale].C);
        Vec3 unit{center[0] - origin[0], center[1] - origin[1], center[2] - origin[2]};
        const double len = norm3(unit);
        const double il = len >= 1e-20 ? 1.0 / len : 0.0;
        for (double& v : unit) v *= il;
        std::size_t ax = std::abs(unit[0]) > std::abs(unit[1]) ? 1 : 0;
        if (std::abs(unit[ax]) > std::abs(unit[2])) ax = 2;
        Vec3 axis{};
        axis[ax] = 1.0;
        Vec3 second = cross(axis, unit);
        const double sl = norm3(second);
        const double isl = sl >= 1e-20 ? 1.0 / sl : 0.0;
        for (double& v : second) v *= isl;
        return {cross(unit, second), second};
    }

    Vec3 gaugeTangentStep(const Options& opt, const State& st, double first, double second)
    {
        const Vec3 origin = targetCoords(st.cams[opt.gaugeAnchor].C);
        const Vec3 center = targetCoords(st.cams[opt.gaugeScale].C);
        const auto b = gaugeBasis(opt, st);
        const Vec3 td{b[0][0] * first + b[1][0] * second, b[0][1] * first + b[1][1] * second,
                      b[0][2] * first + b[1][2] * second};
        const Vec3 dir{center[0] + td[0] - origin[0], center[1] + td[1] - origin[1], center[2] + td[2] - origin[2]};
        const double len = norm3(dir);
        if (len < 1e-20) return targetCoords(origin);
        const double il = 1.0 / len;
        return targetCoords({origin[0] + dir[0] * il * opt.gaugeBaseline, origin[1] + dir[1] * il * opt.gaugeBaseline,
                             origin[2] + dir[2] * il * opt.gaugeBaseline});
    }

    // observationPrimaryTerms (Assembly.cpp:186-342): up to three (block, 2x13 Jacobian) terms: the camera
    // pose (or a rig camera's capture and sensor) and the intrinsic group.
    struct Terms
    {
        int count = 0;
        int blocks[3];
        double J[3][2 * kP];
    };

    Terms primaryTerms(const Options& opt, const Active& a, const State& st, int cam, const Linearization& L,
                       const std::array<Vec3, 2>* basis, const Rig* rig = nullptr)
    {
        Terms t;
        if (rig && !rig->empty())
        {
            // cameraPosePrimaryTerms, rig branch (Assembly.cpp:207-284): the camera's pose Jacobian [rotation |
            // centre] (world-frame increments) chained to its capture (rotation moves the lever arm
            // l = R_rig c_sensor: dC/dw_a = e_a x l) and to its sensor (increments in rig axes: columns R_rig e_a)
            if (rig->bound(cam))
            {
                const Mat3& Rr = st.rig.capR[rig->camCapture[cam]];
                const Vec3& cs = st.rig.senC[rig->camSensor[cam]];
                Vec3 lever{};
                for (int w = 0; w < 3; ++w)
                    for (int r = 0; r < 3; ++r) lever[w] += Rr[w * 3 + r] * cs[r];
                if (a.rigCaptureBlockByCamera[cam] >= 0)
                {
                    double* J = t.J[t.count];
                    std::fill(J, J + 2 * kP, 0.0);
                    for (int r = 0; r < 2; ++r)
                    {
                        const double* src = L.Jpose + r * 6;
                        for (int ax = 0; ax < 3; ++ax)
                        {
                            Vec3 e{};
                            e[ax] = 1.0;
                            const Vec3 d = cross(e, lever);
                            J[r * kP + ax] = src[ax] + src[3] * d[0] + src[4] * d[1] + src[5] * d[2];
                            J[r * kP + 3 + ax] = src[3 + ax];
                        }
                    }
                    t.blocks[t.count++] = a.rigCaptureBlockByCamera[cam];
                }
                if (a.rigSensorBlockByCamera[cam] >= 0)
                {
                    double* J = t.J[t.count];
                    std::fill(J, J + 2 * kP, 0.0);
                    for (int r = 0; r < 2; ++r)
                    {
                        const double* src = L.Jpose + r * 6;
                        for (int ax = 0; ax < 3; ++ax)
                            for (int w = 0; w < 3; ++w)
                            {
                                J[r * kP + ax] += src[w] * Rr[w * 3 + ax];
                                J[r * kP + 3 + ax] += src[3 + w] * Rr[w * 3 + ax];
                            }
                    }
                    t.blocks[t.count++] = a.rigSensorBlockByCamera[cam];
                }
            }
        }
        else if (a.cameraBlock[cam] >= 0)
        {
            double* J = t.J[0];
            std::fill(J, J + 2 * kP, 0.0);
            const bool tangent = usesGaugeTangent(opt, st, cam);
            for (int r = 0; r < 2; ++r)
            {
                const double* src = L.Jpose + r * 6;
                for (int i = 0; i < 3; ++i) J[r * kP + i] = src[i];
                if (tangent)
                {
                    for (int q = 0; q < 2; ++q)
                    {
                        const Vec3 wb = targetCoords((*basis)[q]);
                        for (int ax = 0; ax < 3; ++ax) J[r * kP + 3 + q] += src[3 + ax] * wb[ax];
                    }
                }
                else
                {
                    for (int i = 0; i < 3; ++i) J[r * kP + 3 + i] = src[3 + i];
                }
            }
            t.blocks[0] = a.cameraBlock[cam];
            t.count = 1;
        }
        if (a.intrinsicBlockByCamera[cam] >= 0)
        {
            t.blocks[t.count] = a.intrinsicBlockByCamera[cam];
            std::copy(L.Jint, L.Jint + 2 * kP, t.J[t.count]);
            ++t.count;
        }
        return t;
    }

    // Dense symmetric matrix over padded primary blocks, stored as 13x13 blocks keyed by (row, col) block pairs.
    struct Reduced
    {
        int P = 0;
        std::vector<double> diag;     // P * 169
        std::vector<double> off;      // slots * 169, (row < col)
        std::vector<int> offIndex;    // P*P lookup -> slot (dense; P is small for the oracle)
        std::vector<std::pair<int, int>> offBlocks;
        std::vector<double> directRhs;  // P * 13 (b_c = -sum w J^T r, priors included)
        std::vector<double> schurRhs;   // P * 13 (-sum V b_p)
        std::vector<double> gradient;   // P * 13 (reduced gradient g, solve S d = -g)
        double cost = 0.0;
        bool singular = false;
    };

    struct PointBackSub
    {
        bool valid = false;
        Mat3 inv{};
        Vec3 rhs{};
        std::vector<int> blocks;
        std::vector<std::array<double, kP * kE>> cross;
    };

    struct Oracle
    {
        Problem pb;
        Options opt;
        Active act;
        State init;
        State state;
        bool prepared = false;
        std::string error;
        std::vector<PointBackSub> backSub;
        std::vector<std::pair<int, int>> offBlocks;  // reduced pattern (row < col)
        std::vector<int> offLookup;
        std::vector<std::size_t> trackBounds;
        int threads = 1;
        pb_solve_report report{};
    };

    void prepareTopology(Oracle& o)
    {
        const Active& a = o.act;
        const int P = a.primaryBlockCount;
        o.offBlocks.clear();
        o.offLookup.assign(static_cast<std::size_t>(P) * P, -1);
        std::vector<char> conn(static_cast<std::size_t>(P) * P, 0);
        auto connect = [&](int l, int r)
        {
            if (l < 0 || r < 0 || l == r) return;
            conn[static_cast<std::size_t>(std::min(l, r)) * P + std::max(l, r)] = 1;
        };
        std::vector<int> blocks;
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
        {
            if (!a.activeTrack[t]) continue;
            blocks.clear();
            for (const Observation& ob : o.pb.tracks[t].obs)
            {
                if (!observationUsable(ob, o.pb.cams.size())) continue;
                // the primary blocks of the observation: camera pose (or rig capture and sensor), intrinsics
                const int bs[3] = {o.pb.rig.empty() ? a.cameraBlock[ob.cam] : a.rigCaptureBlockByCamera[ob.cam],
                                   o.pb.rig.empty() ? -1 : a.rigSensorBlockByCamera[ob.cam],
                                   a.intrinsicBlockByCamera[ob.cam]};
                for (int i = 0; i < 3; ++i)
                    for (int j = i + 1; j < 3; ++j) connect(bs[i], bs[j]);
                if (a.trackBlock[t] >= 0)
                    for (int b : bs)
                        if (b >= 0) blocks.push_back(b);
            }
            if (a.trackBlock[t] < 0) continue;
            std::sort(blocks.begin(), blocks.end());
            blocks.erase(std::unique(blocks.begin(), blocks.end()), blocks.end());
            for (std::size_t i = 0; i < blocks.size(); ++i)
                for (std::size_t j = i + 1; j < blocks.size(); ++j) connect(blocks[i], blocks[j]);
        }
        for (int r = 0; r < P; ++r)
            for (int c = r + 1; c < P; ++c)
                if (conn[static_cast<std::size_t>(r) * P + c])
                {
                    o.offLookup[static_cast<std::size_t>(r) * P + c] = static_cast<int>(o.offBlocks.size());
                    o.offBlocks.emplace_back(r, c);
                }
        // contiguous track ranges balanced by usable observation counts (ReferenceSchur.cpp:538-566)
        const int T = std::max(1, std::min<int>(o.threads, static_cast<int>(o.pb.tracks.size())));
        std::vector<std::size_t> prefix(o.pb.tracks.size() + 1, 0);
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
            prefix[t + 1] = prefix[t] + (a.activeTrack[t] ? std::max<std::size_t>(1, o.pb.tracks[t].obs.size()) : 1);
        o.trackBounds.assign(T + 1, o.pb.tracks.size());
        o.trackBounds[0] = 0;
        for (int i = 1; i < T; ++i)
        {
            const std::size_t target = prefix.back() * i / T;
            o.trackBounds[i] = std::lower_bound(prefix.begin(), prefix.end(), target) - prefix.begin();
        }
    }

    // 3x3 inverse with singularity test. Unpinned choice for plamatrix::tryInverse3x3RowMajor(m, inv, 1e-15):
    // fail when |det| <= 1e-15 * max|m_ij|^3 or when anything is non-finite.
    bool inverse3(const Mat3& m, Mat3* inv)
    {
        double mx = 0.0;
        for (double v : m)
        {
            if (!std::isfinite(v)) return false;
            mx = std::max(mx, std::abs(v));
        }
        const double c00 = m[4] * m[8] - m[5] * m[7];
        const double c01 = m[5] * m[6] - m[3] * m[8];
        const double c02 = m[3] * m[7] - m[4] * m[6];
        const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
        if (!std::isfinite(det) || !(std::abs(det) > 1e-15 * mx * mx * mx)) return false;
        const double id = 1.0 / det;
        (*inv)[0] = c00 * id;
        (*inv)[1] = (m[2] * m[7] - m[1] * m[8]) * id;
        (*inv)[2] = (m[1] * m[5] - m[2] * m[4]) * id;
        (*inv)[3] = c01 * id;
        (*inv)[4] = (m[0] * m[8] - m[2] * m[6]) * id;
        (*inv)[5] = (m[2] * m[3] - m[0] * m[5]) * id;
        (*inv)[6] = c02 * id;
        (*inv)[7] = (m[1] * m[6] - m[0] * m[7]) * id;
        (*inv)[8] = (m[0] * m[4] - m[1] * m[3]) * id;
        for (double v : *inv)
            if (!std::isfinite(v)) return false;
        return true;
    }

    // buildReferenceReducedNormalEquations (ReferenceSchur.cpp:675-980).
    Reduced buildReduced(Oracle& o, const State& st, double damping)
    {
        const Active& a = o.act;
        const int P = a.primaryBlockCount;
        const Options& opt = o.opt;
        const bool refParam = opt.referenceParam();
        const std::vector<Camera> eff = effectiveCameras(o.pb, a, st);
        const bool tangent = opt.gaugeScale >= 0 && opt.gaugeAnchor >= 0 && opt.gaugeBaseline > 0.0;
        std::array<Vec3, 2> basis{};
        if (tangent) basis = gaugeBasis(opt, st);
        const int T = static_cast<int>(o.trackBounds.size()) - 1;
        const std::size_t nOff = o.offBlocks.size();
        const std::size_t diagDim = static_cast<std::size_t>(P) * kP * kP;
        const std::size_t vecDim = static_cast<std::size_t>(P) * kP;
        const std::size_t offDim = nOff * kP * kP;
        // worker-exclusive off-diagonal slots while affordable, shared atomic slots otherwise
        const std::size_t budget = 1ull << 30;
        const bool ownSlots = offDim * sizeof(double) * T <= budget;
        std::vector<std::vector<double>> pDiag(T, std::vector<double>(diagDim, 0.0));
        std::vector<std::vector<double>> pDirect(T, std::vector<double>(vecDim, 0.0));
        std::vector<std::vector<double>> pSchur(T, std::vector<double>(vecDim, 0.0));
        std::vector<std::vector<double>> pOff(ownSlots ? T : 1, std::vector<double>(offDim, 0.0));
        std::vector<double> pCost(T, 0.0);
        std::vector<char> pSingular(T, 0), pInvalid(T, 0);
        o.backSub.assign(o.pb.tracks.size(), PointBackSub{});
        const SlotMask noSlots{};

#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int th = 0; th < T; ++th)
        {
            double* diag = pDiag[th].data();
            double* direct = pDirect[th].data();
            double* schur = pSchur[th].data();
            double* off = pOff[ownSlots ? th : 0].data();
            auto hadd = [&](int rb, int cb, int row, int col, double v)
            {
                if (rb == cb)
                {
                    diag[static_cast<std::size_t>(rb) * kP * kP + row * kP + col] += v;
                    return;
                }
                const bool ordered = rb < cb;
                const int r0 = ordered ? rb : cb, c0 = ordered ? cb : rb;
                const int slot = o.offLookup[static_cast<std::size_t>(r0) * P + c0];
                const int sr = ordered ? row : col, sc = ordered ? col : row;
                double& e = off[static_cast<std::size_t>(slot) * kP * kP + sr * kP + sc];
                if (ownSlots)
                    e += v;
                else
                    std::atomic_ref<double>(e).fetch_add(v, std::memory_order_relaxed);
            };
            double cost = 0.0;
            std::vector<int> pblocks;
            std::vector<std::array<double, kP * kE>> cross;
            for (std::size_t t = o.trackBounds[th]; t < o.trackBounds[th + 1]; ++t)
            {
                if (!a.activeTrack[t]) continue;
                const bool elim = a.trackBlock[t] >= 0;
                Mat3 Hpp{};
                Vec3 bp{};
                pblocks.clear();
                cross.clear();
                const Vec3& X = st.points[t];
                for (const Observation& ob : o.pb.tracks[t].obs)
                {
                    if (!observationUsable(ob, st.cams.size())) continue;
                    const int g = a.intrinsicBlockByCamera[ob.cam] >= 0 ? a.calibrationGroupByCamera[ob.cam] : -1;
                    const SlotMask& act = g >= 0 ? st.groups[g].enabled : noSlots;
                    Linearization L;
                    if (!linearize(eff[ob.cam], X, ob, opt.o.robust_loss, opt.o.robust_loss_scale, refParam, act,
                                   &L))
                    {
                        pInvalid[th] = 1;
                        continue;
                    }
                    cost += L.cost;
                    const Terms terms = primaryTerms(opt, a, st, ob.cam, L, tangent ? &basis : nullptr, &o.pb.rig);
                    // accumulateObservationPrimaryTerms (ReferenceSchur.cpp:140-193)
                    int act_idx[3][kP];
                    int act_n[3] = {0, 0, 0};
                    for (int k = 0; k < terms.count; ++k)
                    {
                        for (int q = 0; q < kP; ++q)
                        {
                            const double jx = terms.J[k][q], jy = terms.J[k][kP + q];
                            if (jx == 0.0 && jy == 0.0) continue;
                            act_idx[k][act_n[k]++] = q;
                            const double grad = L.w * (jx * L.r[0] + jy * L.r[1]);
                            direct[static_cast<std::size_t>(terms.blocks[k]) * kP + q] -= grad;
                        }
                    }
                    for (int l = 0; l < terms.count; ++l)
                        for (int r = l; r < terms.count; ++r)
                            for (int i = 0; i < act_n[l]; ++i)
                            {
                                const int row = act_idx[l][i];
                                const double lx = terms.J[l][row], ly = terms.J[l][kP + row];
                                for (int j = 0; j < act_n[r]; ++j)
                                {
                                    const int col = act_idx[r][j];
                                    hadd(terms.blocks[l], terms.blocks[r], row, col,
                                         L.w * (lx * terms.J[r][col] + ly * terms.J[r][kP + col]));
                                }
                            }
                    if (!elim) continue;
                    // accumulatePointTerms (ReferenceSchur.cpp:195-239)
                    for (int c = 0; c < kE; ++c)
                        for (int r = 0; r < 2; ++r)
                        {
                            const double pv = L.Jp[r * kE + c];
                            bp[c] -= L.w * pv * L.r[r];
                            for (int k = 0; k < kE; ++k) Hpp[c * kE + k] += L.w * pv * L.Jp[r * kE + k];
                        }
                    for (int k = 0; k < terms.count; ++k)
                    {
                        auto it = std::find(pblocks.begin(), pblocks.end(), terms.blocks[k]);
                        std::size_t idx;
                        if (it == pblocks.end())
                        {
                            pblocks.push_back(terms.blocks[k]);
                            cross.emplace_back();
                            cross.back().fill(0.0);
                            idx = cross.size() - 1;
                        }
                        else
                        {
                            idx = static_cast<std::size_t>(it - pblocks.begin());
                        }
                        for (int q = 0; q < kP; ++q)
                        {
                            const double jx = terms.J[k][q], jy = terms.J[k][kP + q];
                            if (jx == 0.0 && jy == 0.0) continue;
                            for (int e = 0; e < kE; ++e)
                                cross[idx][q * kE + e] += L.w * (jx * L.Jp[e] + jy * L.Jp[kE + e]);
                        }
                    }
                }
                // point-only survey constraints: laser planes (path A, Assembly.cpp:656-671), then control
                // points (ReferenceSchur.cpp:331-346, accumulatePointOnlyConstraint :241-259); the cost counts
                // for every active track, the normal terms only for an eliminated point; a constraint whose
                // linearisation fails is skipped
                auto addConstraint = [&](const ConstraintLin& c)
                {
                    cost += c.cost;
                    if (!elim) return;
                    for (int col = 0; col < kE; ++col)
                        for (int r = 0; r < c.rows; ++r)
                        {
                            const double pv = c.J[r * kE + col];
                            bp[col] -= c.w * pv * c.r[r];
                            for (int k = 0; k < kE; ++k) Hpp[col * kE + k] += c.w * pv * c.J[r * kE + k];
                        }
                };
                if (opt.laserPlanes)
                    for (const LaserPlane& k : o.pb.tracks[t].planes)
                    {
                        ConstraintLin c;
                        if (linearizePlane(k, X, opt.o, &c)) addConstraint(c);
                    }
                if (opt.controlPoints)
                    for (const ControlPoint& k : o.pb.tracks[t].controls)
                    {
                        ConstraintLin c;
                        if (linearizeControl(k, X, opt.o, &c)) addConstraint(c);
                    }
                if (!elim) continue;
                // eliminatePoint (ReferenceSchur.cpp:350-490); damping form Hpp_ii *= (1 + lambda)
                Mat3 damped = Hpp;
                for (int i = 0; i < kE; ++i) damped[i * kE + i] *= 1.0 + damping;
                Mat3 inv{};
                PointBackSub& bs = o.backSub[t];
                if (!inverse3(damped, &inv))
                {
                    bs.valid = false;
                    pSingular[th] = 1;
                    continue;
                }
                const std::size_t nb = pblocks.size();
                std::vector<std::array<double, kP * kE>> red(nb);
                std::vector<std::array<char, kP>> activeRow(nb);
                for (std::size_t b = 0; b < nb; ++b)
                {
                    red[b].fill(0.0);
                    activeRow[b].fill(0);
                    for (int row = 0; row < kP; ++row)
                    {
                        const double* cr = &cross[b][row * kE];
                        activeRow[b][row] = cr[0] != 0.0 || cr[1] != 0.0 || cr[2] != 0.0;
                        if (!activeRow[b][row]) continue;
                        for (int c = 0; c < kE; ++c)
                            for (int k = 0; k < kE; ++k) red[b][row * kE + c] += cr[k] * inv[k * kE + c];
                    }
                    for (int row = 0; row < kP; ++row)
                    {
                        if (!activeRow[b][row]) continue;
                        double gr = 0.0;
                        for (int c = 0; c < kE; ++c) gr += red[b][row * kE + c] * bp[c];
                        schur[static_cast<std::size_t>(pblocks[b]) * kP + row] -= gr;
                    }
                }
                for (std::size_t l = 0; l < nb; ++l)
                    for (std::size_t r = l; r < nb; ++r)
                        for (int row = 0; row < kP; ++row)
                        {
                            if (!activeRow[l][row]) continue;
                            for (int col = 0; col < kP; ++col)
                            {
                                if (!activeRow[r][col]) continue;
                                double h = 0.0;
                                for (int k = 0; k < kE; ++k) h -= red[l][row * kE + k] * cross[r][col * kE + k];
                                hadd(pblocks[l], pblocks[r], row, col, h);
                            }
                        }
                bs.valid = true;
                bs.inv = inv;
                bs.rhs = bp;
                bs.blocks = pblocks;
                bs.cross = cross;
            }
            pCost[th] = cost;
        }
        for (int th = 0; th < T; ++th)
            if (pInvalid[th]) throw InvalidProjection();

        Reduced R;
        R.P = P;
        R.diag.assign(diagDim, 0.0);
        R.off.assign(offDim, 0.0);
        R.directRhs.assign(vecDim, 0.0);
        R.schurRhs.assign(vecDim, 0.0);
        R.gradient.assign(vecDim, 0.0);
        for (int th = 0; th < T; ++th)
        {
            for (std::size_t i = 0; i < vecDim; ++i)
            {
                R.directRhs[i] += pDirect[th][i];
                R.schurRhs[i] += pSchur[th][i];
            }
            for (std::size_t i = 0; i < diagDim; ++i) R.diag[i] += pDiag[th][i];
            R.cost += pCost[th];
            R.singular = R.singular || pSingular[th];
        }
        for (const auto& slot : pOff)
            for (std::size_t i = 0; i < offDim; ++i) R.off[i] += slot[i];
        for (std::size_t i = 0; i < vecDim; ++i) R.gradient[i] = -R.schurRhs[i] - R.directRhs[i];
        // intrinsic priors (ReferenceSchur.cpp:492-536)
        for (std::size_t gi = 0; gi < st.groups.size(); ++gi)
        {
            const Group& g = st.groups[gi];
            const int block = a.cameraBlockCount + static_cast<int>(gi);
            for (int q = 0; q < kP; ++q)
            {
                if (!g.enabled[q]) continue;
                const double w = g.invSigma[q];
                const double res = w * (g.params[q] - g.prior[q]);
                R.cost += 0.5 * res * res;
                R.diag[static_cast<std::size_t>(block) * kP * kP + q * kP + q] += w * w;
                R.gradient[static_cast<std::size_t>(block) * kP + q] += w * res;
                R.directRhs[static_cast<std::size_t>(block) * kP + q] -= w * res;
            }
        }
        // finalizePrimaryDiagonal(1 + lambda, 1.0). Unpinned choice: every diagonal entry is scaled by
        // (1 + lambda); an entry that is exactly zero (padded or inactive column) becomes 1.0.
        for (int b = 0; b < P; ++b)
            for (int q = 0; q < kP; ++q)
            {
                double& d = R.diag[static_cast<std::size_t>(b) * kP * kP + q * kP + q];
                d = d == 0.0 ? 1.0 : d * (1.0 + damping);
            }
        R.offBlocks = o.offBlocks;
        return R;
    }

    // Dense padded reduced matrix (for small problems and tests).
    void denseReduced(const Reduced& R, std::vector<double>* A)
    {
        const std::size_t n = static_cast<std::size_t>(R.P) * kP;
        A->assign(n * n, 0.0);
        for (int b = 0; b < R.P; ++b)
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                    (*A)[(b * kP + i) * n + b * kP + j] = R.diag[static_cast<std::size_t>(b) * kP * kP + i * kP + j];
        for (std::size_t s = 0; s < R.offBlocks.size(); ++s)
        {
            const int rb = R.offBlocks[s].first, cb = R.offBlocks[s].second;
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                {
                    const double v = R.off[s * kP * kP + i * kP + j];
                    (*A)[(rb * kP + i) * n + cb * kP + j] = v;
                    (*A)[(cb * kP + j) * n + rb * kP + i] = v;
                }
        }
    }

    // Reduced solve S d = -g: Cholesky over the compacted columns (diagonal != 1-only padded columns are
    // decoupled identity rows with zero right-hand side), in reverse Cuthill-McKee block order with an
    // envelope (profile) factorisation. Returns false when the matrix is not numerically SPD.
    bool solveReduced(const Reduced& R, int cameraBlocks, std::vector<double>* step, double* relResidual)
    {
        const int P = R.P;
        const std::size_t n13 = static_cast<std::size_t>(P) * kP;
        step->assign(n13, 0.0);
        // active columns: a column is active when its block row or column carries any non-zero coupling
        std::vector<char> active(n13, 0);
        for (int b = 0; b < P; ++b)
            for (int i = 0; i < kP; ++i)
            {
                bool nz = false;
                for (int j = 0; j < kP && !nz; ++j)
                    nz = (j != i && R.diag[static_cast<std::size_t>(b) * kP * kP + i * kP + j] != 0.0);
                const double d = R.diag[static_cast<std::size_t>(b) * kP * kP + i * kP + i];
                nz = nz || d != 1.0 || R.gradient[static_cast<std::size_t>(b) * kP + i] != 0.0;
                active[b * kP + i] = nz;
            }
        for (std::size_t s = 0; s < R.offBlocks.size(); ++s)
        {
            const int rb = R.offBlocks[s].first, cb = R.offBlocks[s].second;
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                    if (R.off[s * kP * kP + i * kP + j] != 0.0)
                    {
                        active[rb * kP + i] = 1;
                        active[cb * kP + j] = 1;
                    }
        }
        // block adjacency among camera blocks and RCM order; intrinsic blocks (dense rows) are ordered last
        const int nc = std::min(cameraBlocks, P);
        std::vector<std::vector<int>> adj(P);
        for (const auto& pr : R.offBlocks)
        {
            if (pr.first >= nc || pr.second >= nc) continue;
            adj[pr.first].push_back(pr.second);
            adj[pr.second].push_back(pr.first);
        }
        std::vector<int> order;
        order.reserve(P);
        std::vector<char> seen(P, 0);
        std::vector<int> deg(P);
        for (int b = 0; b < P; ++b) deg[b] = static_cast<int>(adj[b].size());
        while (static_cast<int>(order.size()) < nc)
        {
            int start = -1;
            for (int b = 0; b < nc; ++b)
                if (!seen[b] && (start < 0 || deg[b] < deg[start])) start = b;
            std::size_t head = order.size();
            order.push_back(start);
            seen[start] = 1;
            while (head < order.size())
            {
                const int u = order[head++];
                std::vector<int> nb;
                for (int v : adj[u])
                    if (!seen[v]) nb.push_back(v);
                std::sort(nb.begin(), nb.end(), [&](int x, int y) { return deg[x] < deg[y] || (deg[x] == deg[y] && x < y); });
                for (int v : nb)
                {
                    seen[v] = 1;
                    order.push_back(v);
                }
            }
        }
        std::reverse(order.begin(), order.end());
        for (int b = nc; b < P; ++b) order.push_back(b);
        // compact columns in RCM block order
        std::vector<std::size_t> colOf(n13, SIZE_MAX);
        std::vector<std::size_t> var;
        for (int i = 0; i < P; ++i)
        {
            const int b = order[i];
            for (int q = 0; q < kP; ++q)
                if (active[b * kP + q])
                {
                    colOf[b * kP + q] = var.size();
                    var.push_back(static_cast<std::size_t>(b) * kP + q);
                }
        }
        const std::size_t n = var.size();
        if (n == 0) return true;
        // envelope: first column per row
        std::vector<std::size_t> first(n);
        for (std::size_t i = 0; i < n; ++i) first[i] = i;
        auto touch = [&](std::size_t r, std::size_t c)
        {
            if (r < c) std::swap(r, c);
            first[r] = std::min(first[r], c);
        };
        for (int b = 0; b < P; ++b)
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                {
                    const std::size_t ci = colOf[b * kP + i], cj = colOf[b * kP + j];
                    if (ci != SIZE_MAX && cj != SIZE_MAX && R.diag[static_cast<std::size_t>(b) * kP * kP + i * kP + j] != 0.0)
                        touch(ci, cj);
                }
        for (std::size_t s = 0; s < R.offBlocks.size(); ++s)
        {
            const int rb = R.offBlocks[s].first, cb = R.offBlocks[s].second;
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                {
                    const std::size_t ci = colOf[rb * kP + i], cj = colOf[cb * kP + j];
                    if (ci != SIZE_MAX && cj != SIZE_MAX && R.off[s * kP * kP + i * kP + j] != 0.0) touch(ci, cj);
                }
        }
        std::vector<std::size_t> rowStart(n + 1, 0);
        for (std::size_t i = 0; i < n; ++i) rowStart[i + 1] = rowStart[i] + (i - first[i] + 1);
        std::vector<double> env(rowStart[n], 0.0);
        auto at = [&](std::size_t r, std::size_t c) -> double& { return env[rowStart[r] + (c - first[r])]; };
        for (int b = 0; b < P; ++b)
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j <= kP - 1; ++j)
                {
                    const std::size_t ci = colOf[b * kP + i], cj = colOf[b * kP + j];
                    const double v = R.diag[static_cast<std::size_t>(b) * kP * kP + i * kP + j];
                    // zeros are not in the envelope (touch() registered nonzeros only) and need no store
                    if (ci == SIZE_MAX || cj == SIZE_MAX || ci < cj || v == 0.0) continue;
                    at(ci, cj) = v;
                }
        for (std::size_t s = 0; s < R.offBlocks.size(); ++s)
        {
            const int rb = R.offBlocks[s].first, cb = R.offBlocks[s].second;
            for (int i = 0; i < kP; ++i)
                for (int j = 0; j < kP; ++j)
                {
                    std::size_t ci = colOf[rb * kP + i], cj = colOf[cb * kP + j];
                    const double v = R.off[s * kP * kP + i * kP + j];
                    if (ci == SIZE_MAX || cj == SIZE_MAX || v == 0.0) continue;
                    if (ci < cj) std::swap(ci, cj);
                    at(ci, cj) = v;
                }
        }
        std::vector<double> A0 = env;  // keep for the residual check
        // envelope Cholesky L L^T
        for (std::size_t i = 0; i < n; ++i)
        {
            const std::size_t fi = first[i];
            double* Li = &env[rowStart[i]];
            for (std::size_t j = fi; j < i; ++j)
            {
                const std::size_t fj = first[j];
                const std::size_t k0 = std::max(fi, fj);
                const double* Lj = &env[rowStart[j]];
                double s = Li[j - fi];
                for (std::size_t k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
                Li[j - fi] = s / Lj[j - fj];
            }
            double d = Li[i - fi];
            for (std::size_t k = fi; k < i; ++k) d -= Li[k - fi] * Li[k - fi];
            if (!(d > 0.0) || !std::isfinite(d)) return false;
            Li[i - fi] = std::sqrt(d);
        }
        std::vector<double> b(n), y(n);
        for (std::size_t i = 0; i < n; ++i) b[i] = -R.gradient[var[i]];
        for (std::size_t i = 0; i < n; ++i)
        {
            double s = b[i];
            const std::size_t fi = first[i];
            for (std::size_t k = fi; k < i; ++k) s -= env[rowStart[i] + k - fi] * y[k];
            y[i] = s / env[rowStart[i] + i - fi];
        }
        for (std::size_t ii = n; ii-- > 0;)
        {
            y[ii] /= env[rowStart[ii] + ii - first[ii]];
            const double yi = y[ii];
            const std::size_t fi = first[ii];
            for (std::size_t k = fi; k < ii; ++k) y[k] -= env[rowStart[ii] + k - fi] * yi;
        }
        // residual check ||A x - b|| / ||b|| using the lower envelope of A
        std::vector<double> ax(n, 0.0);
        for (std::size_t i = 0; i < n; ++i)
        {
            const std::size_t fi = first[i];
            for (std::size_t k = fi; k <= i; ++k)
            {
                const double v = A0[rowStart[i] + k - fi];
                ax[i] += v * y[k];
                if (k != i) ax[k] += v * y[i];
            }
        }
        double rn = 0.0, bn = 0.0;
        for (std::size_t i = 0; i < n; ++i)
        {
            rn += (ax[i] - b[i]) * (ax[i] - b[i]);
            bn += b[i] * b[i];
        }
        *relResidual = bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
        for (std::size_t i = 0; i < n; ++i) (*step)[var[i]] = y[i];
        return true;
    }

    // recoverReferencePointSteps (ReferenceSchur.cpp:982-1082).
    bool backSubstitute(const Oracle& o, const std::vector<double>& primary, const Reduced& R,
                        std::vector<double>* elim, double* dd)
    {
        const Active& a = o.act;
        elim->assign(static_cast<std::size_t>(a.trackBlockCount) * kE, 0.0);
        double dec = 0.0;
        for (std::size_t i = 0; i < primary.size(); ++i) dec += primary[i] * R.directRhs[i];
        bool ok = true;
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
        {
            const int pbk = a.trackBlock[t];
            if (!a.activeTrack[t] || pbk < 0) continue;
            const PointBackSub& p = o.backSub[t];
            if (!p.valid)
            {
                ok = false;
                continue;
            }
            Vec3 cr = p.rhs;
            for (std::size_t b = 0; b < p.blocks.size(); ++b)
            {
                const std::size_t off = static_cast<std::size_t>(p.blocks[b]) * kP;
                for (int e = 0; e < kE; ++e)
                    for (int q = 0; q < kP; ++q) cr[e] -= p.cross[b][q * kE + e] * primary[off + q];
            }
            for (int r = 0; r < kE; ++r)
            {
                double v = 0.0;
                for (int c = 0; c < kE; ++c) v += p.inv[r * kE + c] * cr[c];
                (*elim)[static_cast<std::size_t>(pbk) * kE + r] = v;
                dec += p.rhs[r] * v;
            }
        }
        *dd = dec;
        return ok;
    }

    struct StepRms
    {
        double global = 0.0, point = 0.0;
    };

    // referenceStepRms (BundleAdjustPlaMatrix.cpp:77-158).
    StepRms stepRms(const Oracle& o, const State& st, const std::vector<double>& primary,
                    const std::vector<double>& elim)
    {
        const Active& a = o.act;
        double gs = 0.0, ps = 0.0;
        std::size_t gd = 0, pd = 0;
        for (int b = 0; b < a.cameraBlockCount; ++b)
        {
            for (int i = 0; i < 6; ++i) gs += primary[b * kP + i] * primary[b * kP + i];
            gd += 6;
        }
        for (std::size_t gi = 0; gi < st.groups.size(); ++gi)
        {
            const std::size_t off = (a.cameraBlockCount + gi) * kP;
            for (int q = 0; q < kP; ++q)
                if (st.groups[gi].enabled[q])
                {
                    gs += primary[off + q] * primary[off + q];
                    ++gd;
                }
        }
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
        {
            const int b = a.trackBlock[t];
            if (b < 0) continue;
            for (int e = 0; e < kE; ++e) ps += elim[b * kE + e] * elim[b * kE + e];
            pd += 3;
        }
        StepRms r;
        r.global = gd ? std::sqrt(gs / gd) : 0.0;
        r.point = pd ? std::sqrt(ps / pd) : 0.0;
        return r;
    }

    // applyStep (Assembly.cpp:1111-1247) for frame cameras; intrinsic clamp (Model.cpp:248-269).
    void applyStep(const Oracle& o, const std::vector<double>& primary, const std::vector<double>& elim,
                   State* st, double alpha)
    {
        const Active& a = o.act;
        const Options& opt = o.opt;
        std::vector<double> sp = primary;
        if (alpha != 1.0)
            for (double& v : sp) v *= alpha;
        const State before = *st;
        if (!o.pb.rig.empty())
        {
            // rig branch (Assembly.cpp:1152-1208): applyPoseDelta on every refined capture and sensor (the
            // world-frame / rig-frame left retraction of the rig Jacobians), then the cameras are composed again
            for (std::size_t k = 0; k < a.rigCaptureBlock.size(); ++k)
                if (a.rigCaptureBlock[k] >= 0)
                    retractExp(&st->rig.capR[k], &st->rig.capC[k],
                               sp.data() + static_cast<std::size_t>(a.rigCaptureBlock[k]) * kP);
            for (std::size_t k = 0; k < a.rigSensorBlock.size(); ++k)
                if (a.rigSensorBlock[k] >= 0)
                    retractExp(&st->rig.senR[k], &st->rig.senC[k],
                               sp.data() + static_cast<std::size_t>(a.rigSensorBlock[k]) * kP);
            composeRig(o.pb.rig, st->rig, &st->cams);
        }
        for (std::size_t c = 0; c < st->cams.size() && o.pb.rig.empty(); ++c)
        {
            const int b = a.cameraBlock[c];
            if (b < 0) continue;
            const double* d = sp.data() + static_cast<std::size_t>(b) * kP;
            Camera& cam = st->cams[c];
            if (!opt.referenceParam())
            {
                retractExp(&cam.R, &cam.C, d);
            }
            else if (usesGaugeTangent(opt, before, static_cast<int>(c)))
            {
                const Vec3 center = gaugeTangentStep(opt, *st, d[3], d[4]);
                const double rot[6] = {d[0], d[1], d[2], 0.0, 0.0, 0.0};
                retractReference(&cam.R, &cam.C, rot);
                cam.C = center;
            }
            else
            {
                retractReference(&cam.R, &cam.C, d);
            }
        }
        for (std::size_t gi = 0; gi < st->groups.size(); ++gi)
        {
            Group& g = st->groups[gi];
            const std::size_t off = (a.cameraBlockCount + gi) * kP;
            for (int q = 0; q < kP; ++q)
                if (g.enabled[q]) g.params[q] = std::clamp(g.params[q] + sp[off + q], g.lower[q], g.upper[q]);
        }
        for (std::size_t t = 0; t < st->points.size(); ++t)
        {
            const int b = a.trackBlock[t];
            if (b < 0) continue;
            const double s = opt.referenceParam() ? pointScale(st->points[t]) : 1.0;
            for (int e = 0; e < kE; ++e) st->points[t][e] += s * alpha * elim[static_cast<std::size_t>(b) * kE + e];
        }
    }

    // evaluateObjective (Assembly.cpp:1010-1019) + intrinsic priors; invalid projection -> +inf.
    double evaluateObjective(const Oracle& o, const State& st)
    {
        const Active& a = o.act;
        const std::vector<Camera> eff = effectiveCameras(o.pb, a, st);
        const int T = static_cast<int>(o.trackBounds.size()) - 1;
        std::vector<double> part(T, 0.0);
        std::vector<char> bad(T, 0);
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int th = 0; th < T; ++th)
        {
            double c = 0.0;
            for (std::size_t t = o.trackBounds[th]; t < o.trackBounds[th + 1] && !bad[th]; ++t)
            {
                if (!a.activeTrack[t]) continue;
                for (const Observation& ob : o.pb.tracks[t].obs)
                {
                    if (!observationUsable(ob, st.cams.size())) continue;
                    double oc = 0.0;
                    if (!observationCost(eff[ob.cam], st.points[t], ob, o.opt.o.robust_loss,
                                         o.opt.o.robust_loss_scale, &oc))
                    {
                        bad[th] = 1;
                        break;
                    }
                    c += oc;
                }
                if (bad[th]) break;
                // survey constraint costs (Assembly.cpp:656-690)
                if (o.opt.laserPlanes)
                    for (const LaserPlane& k : o.pb.tracks[t].planes)
                    {
                        ConstraintLin cl;
                        if (linearizePlane(k, st.points[t], o.opt.o, &cl)) c += cl.cost;
                    }
                if (o.opt.controlPoints)
                    for (const ControlPoint& k : o.pb.tracks[t].controls)
                    {
                        ConstraintLin cl;
                        if (linearizeControl(k, st.points[t], o.opt.o, &cl)) c += cl.cost;
                    }
            }
            part[th] = c;
        }
        double cost = 0.0;
        for (int th = 0; th < T; ++th)
        {
            if (bad[th]) return kInf;
            cost += part[th];
        }
        for (const Group& g : st.groups)
            for (int q = 0; q < kP; ++q)
                if (g.enabled[q])
                {
                    const double r = g.invSigma[q] * (g.params[q] - g.prior[q]);
                    cost += 0.5 * r * r;
                }
        return cost;
    }

    void setMessage(pb_solve_report* r, const std::string& m)
    {
        std::snprintf(r->message, sizeof(r->message), "%s", m.c_str());
    }

    // canUseReferenceOnlineSchur (ReferenceSchur.cpp:667-673) for frame cameras without rigs, laser ranges,
    // scale bars, pose priors or a camera plane.
    bool onlineSchur(const Oracle& o)
    {
        return o.opt.o.use_online_schur != 0 && o.act.primaryBlockCount > 0 && !o.opt.laserPlanes && !o.opt.rig;
    }

    void finishReport(Oracle& o, const State& st, double current, bool converged, bool cancelled, bool failed,
                      int accepted, int rejected, int iterations, std::chrono::steady_clock::time_point t0)
    {
        pb_solve_report& rep = o.report;
        rep.solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep.accepted_steps = accepted;
        rep.rejected_steps = rejected;
        rep.iterations = iterations;
        rep.final_cost = current;
        rep.online_schur_used = onlineSchur(o) ? 1 : 0;
        if (cancelled)
        {
            o.state = o.init;
            rep.solve_status = PB_SOLVE_CANCELLED;
            rep.solution_usable = 0;
            setMessage(&rep, "cancelled; no intermediate solution published");
            return;
        }
        if (failed)
        {
            o.state = o.init;
            return;
        }
        rep.converged = converged;
        rep.solve_status = converged ? PB_SOLVE_SUCCESS : PB_SOLVE_NO_CONVERGENCE;
        rep.solution_usable = std::isfinite(current);
        setMessage(&rep, converged ? "converged" : "iteration budget or line search exhausted");
        o.state = st;
    }

    // optimizePointsWithPlaMatrix, assembled branch (path A, BundleAdjustPlaMatrix.cpp:239-556 with the online
    // Schur off): one linearisation per outer step; after a rejected step the same equations are solved again
    // with more damping; the directional derivative comes from an objective probe at alpha = 1e-4 (:473-486);
    // any decrease is accepted (:517-519); a small step ends the solve without being applied (:451-465).
    // Unpinned choices (PlaMatrix's solveDampedSchurComplement is external): the damped system is the online
    // path's (Hpp and reduced diagonals times 1 + lambda); a singular damped point block is a linear-solver
    // exception, which this path rethrows (:357-364); a reduced system that is not positive definite counts as a
    // solve that did not converge (more damping, the same equations).
    void solveAssembled(Oracle& o, const volatile int32_t* cancel)
    {
        pb_solve_report& rep = o.report;
        const Active& a = o.act;
        const auto t0 = std::chrono::steady_clock::now();
        State st = o.init;
        double current = std::numeric_limits<double>::quiet_NaN();
        double damping = 0.0;
        int accepted = 0, rejected = 0, iterations = 0;
        bool converged = a.primaryBlockCount == 0 && a.trackBlockCount == 0;
        bool cancelled = false, failed = false;
        auto isCancelled = [&]() { return cancel && *cancel != 0; };
        auto raise = [&]() { damping = damping == 0.0 ? 1e-3 : damping * 10.0; };
        const double tol = o.opt.o.step_rms_tolerance > 0.0 ? o.opt.o.step_rms_tolerance : 1e-6;
        rep.linear_tolerance = 1e-12;
        try
        {
            if (converged)
            {
                current = evaluateObjective(o, st);
                rep.initial_cost = current;
                ++rep.objective_evaluations;
            }
            const int budget = std::max(1, o.opt.o.max_iterations);
            while (iterations < budget && !converged)
            {
                if (isCancelled())
                {
                    cancelled = true;
                    break;
                }
                // buildNormalEquations at the current state (:249-269)
                const auto ta = std::chrono::steady_clock::now();
                Reduced R = buildReduced(o, st, damping);
                current = R.cost;
                rep.assembly_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - ta).count();
                ++rep.linearizations;
                ++rep.objective_evaluations;
                if (rep.linearizations == 1) rep.initial_cost = current;
                bool relinearize = false;
                bool first = true;
                while (iterations < budget && !converged && !relinearize)
                {
                    if (isCancelled())
                    {
                        cancelled = true;
                        break;
                    }
                    // the same linearisation point, damped with the current damping
                    if (!first) R = buildReduced(o, st, damping);
                    first = false;
                    if (R.singular) throw std::runtime_error("singular damped point block");
                    std::vector<double> primary;
                    double rel = 0.0;
                    const auto tl = std::chrono::steady_clock::now();
                    const bool spd = solveReduced(R, a.cameraBlockCount, &primary, &rel);
                    rep.linear_solve_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();
                    ++iterations;
                    // a reduced system that is not positive definite counts as a solve that did not converge
                    if (!spd || !(rel <= 1e-12))
                    {
                        ++rejected;
                        raise();
                        continue;
                    }
                    std::vector<double> elim;
                    double ddUnused = 0.0;
                    if (!backSubstitute(o, primary, R, &elim, &ddUnused))
                        throw std::runtime_error("singular damped point block");
                    const StepRms rms = stepRms(o, st, primary, elim);
                    if (rms.global <= tol && rms.point <= tol)
                    {
                        converged = true;
                        continue;
                    }
                    const auto tobj = std::chrono::steady_clock::now();
                    State trial = st;
                    applyStep(o, primary, elim, &trial, 1e-4);
                    const double probe = evaluateObjective(o, trial);
                    ++rep.objective_evaluations;
                    double dd = 0.0;
                    if (std::isfinite(probe) && probe < current) dd = (current - probe) / 1e-4;
                    double alpha = 1.0, cand = kInf;
                    for (int ls = 0; ls < o.opt.o.line_search_steps; ++ls)
                    {
                        trial = st;
                        applyStep(o, primary, elim, &trial, alpha);
                        cand = evaluateObjective(o, trial);
                        ++rep.objective_evaluations;
                        const double limit = dd > 0.0 ? current - o.opt.o.armijo_coefficient * alpha * dd : current;
                        if (std::isfinite(cand) && cand <= limit && cand < current) break;
                        alpha *= 0.5;
                    }
                    rep.objective_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - tobj).count();
                    if (std::getenv("PB_ORACLE_TRACE"))
                        std::fprintf(stderr, "oracle A it=%d cost=%.10g cand=%.10g alpha=%g rms_g=%.3e rms_p=%.3e "
                                     "dd=%.4e lambda=%g\n",
                                     iterations, current, cand, alpha, rms.global, rms.point, dd, damping);
                    if (std::isfinite(cand) && cand < current)
                    {
                        st = trial;
                        current = cand;
                        ++accepted;
                        damping /= 10.0;
                        // alpha is the accepted length, or half the last one tried when no length met Armijo
                        converged = alpha * rms.global <= tol && alpha * rms.point <= tol;
                        relinearize = !converged;
                    }
                    else
                    {
                        ++rejected;
                        raise();
                    }
                }
            }
        }
        catch (const std::exception& e)
        {
            failed = true;
            rep.solve_status = PB_SOLVE_NUMERICAL_FAILURE;
            setMessage(&rep, std::string("numerical failure: ") + e.what());
        }
        finishReport(o, st, current, converged, cancelled, failed, accepted, rejected, iterations, t0);
    }

    // optimizePointsWithPlaMatrix, online Schur branch (BundleAdjustPlaMatrix.cpp:162-631).
    void solve(Oracle& o, const volatile int32_t* cancel)
    {
        if (!onlineSchur(o))
        {
            solveAssembled(o, cancel);
            return;
        }
        pb_solve_report& rep = o.report;
        const Active& a = o.act;
        const auto t0 = std::chrono::steady_clock::now();
        State st = o.init;
        double current = std::numeric_limits<double>::quiet_NaN();
        double damping = 0.0;
        int accepted = 0, rejected = 0, iterations = 0;
        bool converged = a.primaryBlockCount == 0 && a.trackBlockCount == 0;
        bool cancelled = false;
        bool failed = false;
        auto isCancelled = [&]() { return cancel && *cancel != 0; };
        try
        {
            if (converged)
            {
                current = evaluateObjective(o, st);
                rep.initial_cost = current;
                ++rep.objective_evaluations;
            }
            const int budget = std::max(1, o.opt.o.max_iterations);
            bool terminate = false;
            while (iterations < budget && !converged && !terminate)
            {
                if (isCancelled())
                {
                    cancelled = true;
                    break;
                }
                bool relinearize = false;
                while (iterations < budget && !converged && !relinearize)
                {
                    if (isCancelled())
                    {
                        cancelled = true;
                        break;
                    }
                    const auto ta = std::chrono::steady_clock::now();
                    Reduced R = buildReduced(o, st, damping);
                    rep.assembly_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - ta).count();
                    ++rep.linearizations;
                    ++rep.objective_evaluations;
                    current = R.cost;
                    if (rep.linearizations == 1) rep.initial_cost = current;
                    if (R.singular)
                    {
                        ++iterations;
                        ++rejected;
                        damping = damping == 0.0 ? 1e-3 : damping * 10.0;
                        continue;
                    }
                    std::vector<double> primary;
                    double rel = 0.0;
                    const auto tl = std::chrono::steady_clock::now();
                    const bool ok = solveReduced(R, o.act.cameraBlockCount, &primary, &rel);
                    rep.linear_solve_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();
                    rep.linear_tolerance = 2e-10;
                    if (!ok)
                    {
                        // retryable "not positive definite" failure
                        ++iterations;
                        ++rejected;
                        damping = damping == 0.0 ? 1e-3 : damping * 10.0;
                        continue;
                    }
                    ++iterations;
                    if (!(rel <= 2e-10))
                    {
                        ++rejected;
                        damping = damping == 0.0 ? 1e-3 : damping * 10.0;
                        continue;
                    }
                    std::vector<double> elim;
                    double dd = 0.0;
                    const auto tb = std::chrono::steady_clock::now();
                    const bool bok = backSubstitute(o, primary, R, &elim, &dd);
                    rep.back_substitution_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
                    if (!bok || !std::isfinite(dd))
                    {
                        ++rejected;
                        damping = damping == 0.0 ? 1e-3 : damping * 10.0;
                        continue;
                    }
                    const StepRms rms = stepRms(o, st, primary, elim);
                    double alpha = 1.0;
                    double cand = kInf;
                    bool armijo = false;
                    State trial;
                    const auto tobj = std::chrono::steady_clock::now();
                    for (int ls = 0; ls < o.opt.o.line_search_steps; ++ls)
                    {
                        trial = st;
                        applyStep(o, primary, elim, &trial, alpha);
                        cand = evaluateObjective(o, trial);
                        ++rep.objective_evaluations;
                        const double limit = current - o.opt.o.armijo_coefficient * alpha * dd;
                        if (std::isfinite(cand) && cand <= limit && cand < current)
                        {
                            armijo = true;
                            break;
                        }
                        alpha *= 0.5;
                    }
                    rep.objective_seconds +=
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - tobj).count();
                    if (std::getenv("PB_ORACLE_TRACE"))
                        std::fprintf(stderr, "oracle it=%d cost=%.10g cand=%.10g alpha=%g rms_g=%.3e rms_p=%.3e dd=%.4e lambda=%g\n",
                                     iterations, current, cand, alpha, rms.global, rms.point, dd, damping);
                    if (std::isfinite(cand) && cand < current && armijo)
                    {
                        st = trial;
                        current = cand;
                        ++accepted;
                        damping /= 10.0;
                        // step-RMS stop (BundleAdjustPlaMatrix.cpp:446-466), threshold 1e-6 unless the options set one
                        const double tol = o.opt.o.step_rms_tolerance > 0.0 ? o.opt.o.step_rms_tolerance : 1e-6;
                        converged = alpha * rms.global <= tol && alpha * rms.point <= tol;
                        relinearize = !converged;
                    }
                    else
                    {
                        ++rejected;
                        terminate = true;
                        relinearize = true;
                    }
                }
            }
        }
        catch (const std::exception& e)
        {
            failed = true;
            rep.solve_status = PB_SOLVE_NUMERICAL_FAILURE;
            setMessage(&rep, std::string("numerical failure: ") + e.what());
        }
        finishReport(o, st, current, converged, cancelled, failed, accepted, rejected, iterations, t0);
    }

    bool prepare(Oracle& o, const pb_solve_options* opts)
    {
        o.opt = Options{};
        o.opt.o = *opts;
        std::memset(&o.report, 0, sizeof(o.report));
        o.report.gauge_anchor_camera = -1;
        o.report.gauge_scale_camera = -1;
        const pb_solve_options& so = *opts;
        if (so.max_iterations <= 0 || !std::isfinite(so.robust_loss_scale) || so.robust_loss_scale <= 0.0 ||
            so.robust_loss < 0 || so.robust_loss > 2 || !(so.armijo_coefficient > 0.0) ||
            !(so.armijo_coefficient < 1.0) || so.line_search_steps <= 0 || so.line_search_steps > 64 ||
            !std::isfinite(so.max_initial_track_rms) || so.max_initial_track_rms < 0.0)
        {
            o.report.solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(&o.report, "invalid options");
            return false;
        }
        if (!o.pb.constraintError.empty() || !o.pb.rigError.empty())
        {
            o.report.solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(&o.report, o.pb.rigError.empty() ? o.pb.constraintError : o.pb.rigError);
            return false;
        }
        o.opt.rig = !o.pb.rig.empty();
        for (const Track& tr : o.pb.tracks)
        {
            o.opt.controlPoints = o.opt.controlPoints || !tr.controls.empty();
            o.opt.laserPlanes = o.opt.laserPlanes || !tr.planes.empty();
        }
        // BundleAdjustValidation.cpp:589-595
        if (o.opt.laserPlanes && (!std::isfinite(so.laser_plane_weight) || so.laser_plane_weight <= 0.0 ||
                                  !std::isfinite(so.laser_huber_delta) || so.laser_huber_delta < 0.0))
        {
            o.report.solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(&o.report, "invalid laser-plane weight or Huber delta");
            return false;
        }
        const std::string gerr = normalizeGauge(o.pb, &o.opt);
        if (!gerr.empty())
        {
            o.report.solve_status = gerr.rfind("unsupported", 0) == 0 ? PB_SOLVE_UNSUPPORTED : PB_SOLVE_INVALID_INPUT;
            setMessage(&o.report, gerr);
            return false;
        }
        o.act = prepareActive(o.pb, o.opt);
        o.init.cams = o.pb.cams;
        o.init.rig = o.pb.rig.poses;
        o.init.points.resize(o.pb.tracks.size());
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t) o.init.points[t] = o.pb.tracks[t].X;
        o.init.groups = initGroups(o.pb, o.opt, o.act);
        o.state = o.init;
        prepareTopology(o);
        const Active& a = o.act;
        pb_solve_report& r = o.report;
        r.observation_count = a.observationCount;
        r.active_track_count = a.activeTrackCount;
        r.rejected_initial_tracks = a.rejectedInitialTracks;
        r.camera_block_count = a.cameraBlockCount;
        r.intrinsic_block_count = a.intrinsicBlockCount;
        r.track_block_count = a.trackBlockCount;
        r.primary_block_count = a.primaryBlockCount;
        r.gauge_anchor_camera = o.opt.gaugeAnchor;
        r.gauge_scale_camera = o.opt.gaugeScale;
        r.gauge_baseline = o.opt.gaugeBaseline;
        if (o.pb.cams.empty() || o.pb.tracks.empty())
        {
            r.solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(&r, "empty cameras or tracks");
            return false;
        }
        if (a.observationCount == 0)
        {
            r.solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(&r, "no usable tracks");
            return false;
        }
        o.prepared = true;
        return true;
    }

    // --- post-solve quality (BundleAdjustQuality.cpp:302-502, solver.cpp:123-171) -------------------
    // strictTrackRms (BundleAdjustQuality.cpp:302-363): whitened weights w / s^2; any weighted observation with an
    // invalid camera or pixel, or whose projection fails, makes the whole track unusable (inf); fewer than two
    // residuals or a single distinct camera also give inf.
    double strictTrackRms(const std::vector<Camera>& cams, const Track& tr, const Vec3& X)
    {
        const double inf = std::numeric_limits<double>::infinity();
        if (!std::isfinite(X[0]) || !std::isfinite(X[1]) || !std::isfinite(X[2])) return inf;
        double sum = 0.0;
        int residuals = 0, first = -1;
        bool second = false;
        for (const Observation& ob : tr.obs)
        {
            const double sc = sanitizedScale(ob);
            const double w = sanitizedWeight(ob) / (sc * sc);
            if (!(w > 0.0)) continue;
            if (ob.cam < 0 || ob.cam >= static_cast<int>(cams.size()) || !std::isfinite(ob.u) || !std::isfinite(ob.v))
                return inf;
            ProjectionOut pr;
            if (!project(cams[ob.cam], X, false, &pr)) return inf;
            const double du = pr.uv[0] - ob.u, dv = pr.uv[1] - ob.v;
            sum += w * (du * du + dv * dv);
            residuals += 2;
            if (first < 0)
                first = ob.cam;
            else if (ob.cam != first)
                second = true;
        }
        if (residuals < 4 || !second) return inf;
        return std::sqrt(sum / residuals);
    }

    // adaptivePointFilterThreshold (BundleAdjustQuality.cpp:373-392)
    double adaptiveThreshold(const std::vector<double>& rms, double floorPx, double factor)
    {
        std::vector<double> finite;
        for (double v : rms)
            if (std::isfinite(v) && v >= 0.0) finite.push_back(v);
        double threshold = std::isfinite(floorPx) && floorPx >= 0.0 ? floorPx : 0.0;
        if (finite.empty() || !std::isfinite(factor) || !(factor > 0.0)) return threshold;
        return std::max(threshold, factor * medianOr(finite, 0.0));
    }
}  // namespace

extern "C" {



struct oracle_ctx
{
    Oracle o;
};

oracle_ctx* oracle_create(const pb_problem_view* v)
{
    auto* ctx = new oracle_ctx();
    Problem& pb = ctx->o.pb;
    pb.cams.resize(v->num_cameras);
    pb.refIntr.resize(v->num_cameras);
    for (int c = 0; c < v->num_cameras; ++c)
    {
        Camera& cam = pb.cams[c];
        std::copy(v->cam_rotation + 9 * c, v->cam_rotation + 9 * c + 9, cam.R.begin());
        std::copy(v->cam_center + 3 * c, v->cam_center + 3 * c + 3, cam.C.begin());
        std::copy(v->cam_intrinsics + kP * c, v->cam_intrinsics + kP * c + kP, cam.p.begin());
        cam.conv = v->cam_convention ? v->cam_convention[c] : PB_BROWN_OPENCV;
        if (v->cam_image_size)
        {
            cam.width = v->cam_image_size[2 * c];
            cam.height = v->cam_image_size[2 * c + 1];
        }
        const double* ref = v->ref_intrinsics ? v->ref_intrinsics + kP * c : v->cam_intrinsics + kP * c;
        std::copy(ref, ref + kP, pb.refIntr[c].begin());
    }
    pb.hasGroupIds = v->cam_group != nullptr;
    if (v->cam_group) pb.groupIds.assign(v->cam_group, v->cam_group + v->num_cameras);
    pb.tracks.resize(v->num_tracks);
    for (int t = 0; t < v->num_tracks; ++t)
    {
        Track& tr = pb.tracks[t];
        std::copy(v->track_point + 3 * t, v->track_point + 3 * t + 3, tr.X.begin());
        for (int64_t i = v->track_offset[t]; i < v->track_offset[t + 1]; ++i)
        {
            Observation ob;
            ob.cam = v->obs_camera[i];
            ob.u = v->obs_uv[2 * i];
            ob.v = v->obs_uv[2 * i + 1];
            ob.weight = v->obs_weight ? v->obs_weight[i] : 1.0;
            ob.scale = v->obs_scale ? v->obs_scale[i] : 1.0;
            tr.obs.push_back(ob);
        }
    }
    if (v->fixed_cameras) pb.fixedCameras.assign(v->fixed_cameras, v->fixed_cameras + v->num_fixed_cameras);
    // camera rig: every camera bound to a capture and a sensor, rotations orthonormal; the bound cameras are
    // composed from it (solver.cpp:282); without calibration group ids each sensor is a group and unbound cameras
    // take group 0 (Conversion.cpp:83-98); a camera whose capture and sensor are both fixed is a fixed camera
    // (Conversion.cpp:120-142)
    if (v->num_rig_captures > 0 || v->num_rig_sensors > 0)
    {
        Rig& rig = pb.rig;
        auto rotationOk = [](const double* R)
        {
            for (int i = 0; i < 9; ++i)
                if (!std::isfinite(R[i])) return false;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    double d = 0.0;
                    for (int k = 0; k < 3; ++k) d += R[k * 3 + i] * R[k * 3 + j];
                    if (std::abs(d - (i == j ? 1.0 : 0.0)) > 1e-6) return false;
                }
            const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                               R[2] * (R[3] * R[7] - R[4] * R[6]);
            return det > 0.0;
        };
        const int nk = std::max(0, v->num_rig_captures), ns = std::max(0, v->num_rig_sensors);
        rig.poses.capR.resize(nk);
        rig.poses.capC.resize(nk);
        rig.capFixed.assign(nk, 0);
        rig.poses.senR.resize(ns);
        rig.poses.senC.resize(ns);
        rig.senFixed.assign(ns, 0);
        for (int k = 0; k < nk; ++k)
        {
            std::copy(v->rig_capture_rotation + 9 * k, v->rig_capture_rotation + 9 * k + 9, rig.poses.capR[k].begin());
            std::copy(v->rig_capture_center + 3 * k, v->rig_capture_center + 3 * k + 3, rig.poses.capC[k].begin());
            rig.capFixed[k] = v->rig_capture_fixed && v->rig_capture_fixed[k] ? 1 : 0;
            if (!rotationOk(rig.poses.capR[k].data()) || !std::isfinite(rig.poses.capC[k][0]) ||
                !std::isfinite(rig.poses.capC[k][1]) || !std::isfinite(rig.poses.capC[k][2]))
                pb.rigError = "rig capture " + std::to_string(k) + " is not a valid rigid transform";
        }
        for (int k = 0; k < ns; ++k)
        {
            std::copy(v->rig_sensor_rotation + 9 * k, v->rig_sensor_rotation + 9 * k + 9, rig.poses.senR[k].begin());
            std::copy(v->rig_sensor_center + 3 * k, v->rig_sensor_center + 3 * k + 3, rig.poses.senC[k].begin());
            rig.senFixed[k] = v->rig_sensor_fixed && v->rig_sensor_fixed[k] ? 1 : 0;
            if (!rotationOk(rig.poses.senR[k].data()) || !std::isfinite(rig.poses.senC[k][0]) ||
                !std::isfinite(rig.poses.senC[k][1]) || !std::isfinite(rig.poses.senC[k][2]))
                pb.rigError = "rig sensor " + std::to_string(k) + " is not a valid rigid transform";
        }
        rig.camCapture.assign(v->num_cameras, -1);
        rig.camSensor.assign(v->num_cameras, -1);
        for (int c = 0; c < v->num_cameras; ++c)
        {
            const int k = v->cam_rig_capture ? v->cam_rig_capture[c] : -1;
            const int q = v->cam_rig_sensor ? v->cam_rig_sensor[c] : -1;
            if (k < 0 || k >= nk || q < 0 || q >= ns)
            {
                pb.rigError = "rig binding of camera " + std::to_string(c) + " is missing or out of range";
                continue;
            }
            rig.camCapture[c] = k;
            rig.camSensor[c] = q;
        }
        if (pb.rigError.empty())
        {
            composeRig(rig, rig.poses, &pb.cams);
            if (!pb.hasGroupIds)
            {
                pb.hasGroupIds = true;
                pb.groupIds.assign(v->num_cameras, 0);
                std::vector<int> keys;
                for (int c = 0; c < v->num_cameras; ++c)
                {
                    auto it = std::find(keys.begin(), keys.end(), rig.camSensor[c]);
                    if (it == keys.end()) it = keys.insert(keys.end(), rig.camSensor[c]);
                    pb.groupIds[c] = static_cast<int>(it - keys.begin());
                }
            }
            for (int c = 0; c < v->num_cameras; ++c)
                if (rig.capFixed[rig.camCapture[c]] && rig.senFixed[rig.camSensor[c]] &&
                    std::find(pb.fixedCameras.begin(), pb.fixedCameras.end(), c) == pb.fixedCameras.end())
                    pb.fixedCameras.push_back(c);
        }
    }
    if (v->fixed_tracks) pb.fixedTracks.assign(v->fixed_tracks, v->fixed_tracks + v->num_fixed_tracks);
    pb.gaugePolicy = v->gauge_policy;
    pb.gaugeAnchor = v->gauge_anchor_camera;
    pb.gaugeScale = v->gauge_scale_camera;
    pb.gaugeBaseline = v->gauge_baseline;
    // survey constraints, validated as validateProblem does (problem.cpp:241-259)
    auto finite3 = [](const double* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); };
    for (int i = 0; i < v->num_control_points; ++i)
    {
        const int t = v->cp_track[i];
        ControlPoint k;
        k.index = i;
        std::copy(v->cp_point + 3 * i, v->cp_point + 3 * i + 3, k.P.begin());
        k.weight = v->cp_weight ? v->cp_weight[i] : 1.0;
        const int kind = v->cp_uncertainty ? v->cp_uncertainty[i] : PB_CONTROL_ISOTROPIC_SIGMA;
        const double sigma = v->cp_sigma ? v->cp_sigma[i] : 1.0;
        const double* m = v->cp_matrix ? v->cp_matrix + 9 * i : nullptr;
        if (t < 0 || t >= v->num_tracks)
        {
            pb.constraintError = "control point track index out of range";
            continue;
        }
        if (!finite3(v->cp_point + 3 * i) || !std::isfinite(k.weight) || k.weight <= 0.0 ||
            !controlWhitening(kind, sigma, m, &k.A))
        {
            pb.constraintError = "track contains an invalid control-point constraint";
            continue;
        }
        pb.tracks[t].controls.push_back(k);
    }
    for (int i = 0; i < v->num_laser_planes; ++i)
    {
        const int t = v->lp_track[i];
        LaserPlane k;
        std::copy(v->lp_point + 3 * i, v->lp_point + 3 * i + 3, k.p.begin());
        std::copy(v->lp_normal + 3 * i, v->lp_normal + 3 * i + 3, k.n.begin());
        k.weight = v->lp_weight ? v->lp_weight[i] : 1.0;
        if (t < 0 || t >= v->num_tracks)
        {
            pb.constraintError = "laser plane track index out of range";
            continue;
        }
        const double nn = std::hypot(k.n[0], k.n[1], k.n[2]);
        if (!finite3(k.p.data()) || !finite3(k.n.data()) || !std::isfinite(nn) || nn <= 0.0 ||
            !std::isfinite(k.weight) || k.weight <= 0.0)
        {
            pb.constraintError = "track contains an invalid laser-plane constraint";
            continue;
        }
        pb.tracks[t].planes.push_back(k);
    }
    return ctx;
}

// Whitening matrix of one control point (makeControlPointWhitening); returns 0 when it is rejected.
int oracle_control_whitening(int kind, double sigma, const double* matrix, double* A)
{
    Mat3 W;
    if (!controlWhitening(kind, sigma, matrix, &W)) return 0;
    std::copy(W.begin(), W.end(), A);
    return 1;
}

void oracle_destroy(oracle_ctx* ctx) { delete ctx; }

void oracle_set_threads(oracle_ctx* ctx, int threads) { ctx->o.threads = std::max(1, threads); }

int oracle_max_threads()
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Full solve. Returns the pb_solve_status.
int oracle_solve(oracle_ctx* ctx, const pb_solve_options* opts, pb_solve_report* report,
                 const volatile int32_t* cancel)
{
    Oracle& o = ctx->o;
    const auto t0 = std::chrono::steady_clock::now();
    if (prepare(o, opts))
    {
        o.report.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        solve(o, cancel);
    }
    else
    {
        o.state = o.init;
    }
    std::snprintf(o.report.device_name, sizeof(o.report.device_name), "cpu-oracle x%d", o.threads);
    if (report) *report = o.report;
    return o.report.solve_status;
}

void oracle_download(oracle_ctx* ctx, double* R, double* C, double* intr, double* X, int32_t* valid)
{
    Oracle& o = ctx->o;
    const State& st = o.state.cams.empty() ? o.init : o.state;
    for (std::size_t c = 0; c < o.pb.cams.size(); ++c)
    {
        const Camera& cam = st.cams.empty() ? o.pb.cams[c] : st.cams[c];
        if (R) std::copy(cam.R.begin(), cam.R.end(), R + 9 * c);
        if (C) std::copy(cam.C.begin(), cam.C.end(), C + 3 * c);
        if (intr)
        {
            // publishIntrinsics (Model.cpp:271-352) without transition priors
            Slots pub = o.pb.cams[c].p;
            if (!st.groups.empty() && o.prepared)
            {
                const Group& g = st.groups[o.act.calibrationGroupByCamera[c]];
                const Slots& ref = o.pb.refIntr[c];
                const double focal = g.enabled[0] ? g.params[0] : o.pb.cams[c].p[0];
                const double aspect = g.enabled[1] ? std::exp(g.params[1]) : std::exp(o.pb.cams[c].p[1]);
                pub[0] = focal;
                pub[1] = std::log(aspect);
                pub[2] = g.enabled[2] ? ref[2] + g.params[2] : o.pb.cams[c].p[2];
                pub[3] = g.enabled[3] ? ref[3] + g.params[3] : o.pb.cams[c].p[3];
                for (int q = 4; q < kP; ++q)
                    if (g.enabled[q]) pub[q] = g.params[q];
            }
            std::copy(pub.begin(), pub.end(), intr + kP * c);
        }
    }
    for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
    {
        const Vec3& p = st.points.empty() ? o.pb.tracks[t].X : st.points[t];
        if (X) std::copy(p.begin(), p.end(), X + 3 * t);
        if (valid) valid[t] = o.prepared ? o.act.activeTrack[t] : 0;
    }
}

void oracle_masks(oracle_ctx* ctx, int32_t* cameraBlock, int32_t* intrBlock, int32_t* groupByCam,
                  int32_t* trackBlock, int32_t* activeTrack)
{
    Oracle& o = ctx->o;
    for (std::size_t c = 0; c < o.pb.cams.size(); ++c)
    {
        if (cameraBlock) cameraBlock[c] = o.prepared ? o.act.cameraBlock[c] : -1;
        if (intrBlock) intrBlock[c] = o.prepared ? o.act.intrinsicBlockByCamera[c] : -1;
        if (groupByCam) groupByCam[c] = o.prepared ? o.act.calibrationGroupByCamera[c] : 0;
    }
    for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
    {
        if (trackBlock) trackBlock[t] = o.prepared ? o.act.trackBlock[t] : -1;
        if (activeTrack) activeTrack[t] = o.prepared ? o.act.activeTrack[t] : 0;
    }
}

// Rig block numbering (rigCaptureBlockByCamera, rigSensorBlockByCamera) of the prepared problem.
void oracle_rig_masks(oracle_ctx* ctx, int32_t* captureBlock, int32_t* sensorBlock)
{
    Oracle& o = ctx->o;
    for (std::size_t c = 0; c < o.pb.cams.size(); ++c)
    {
        const bool rig = o.prepared && !o.act.rigCaptureBlockByCamera.empty();
        if (captureBlock) captureBlock[c] = rig ? o.act.rigCaptureBlockByCamera[c] : -1;
        if (sensorBlock) sensorBlock[c] = rig ? o.act.rigSensorBlockByCamera[c] : -1;
    }
}

// Rig poses of the current state (the input rig before a solve).
void oracle_download_rig(oracle_ctx* ctx, double* capR, double* capC, double* senR, double* senC)
{
    Oracle& o = ctx->o;
    const RigPoses& r = o.state.rig.capR.size() == o.pb.rig.poses.capR.size() && o.prepared ? o.state.rig : o.pb.rig.poses;
    for (std::size_t k = 0; k < r.capR.size(); ++k)
    {
        if (capR) std::copy(r.capR[k].begin(), r.capR[k].end(), capR + 9 * k);
        if (capC) std::copy(r.capC[k].begin(), r.capC[k].end(), capC + 3 * k);
    }
    for (std::size_t k = 0; k < r.senR.size(); ++k)
    {
        if (senR) std::copy(r.senR[k].begin(), r.senR[k].end(), senR + 9 * k);
        if (senC) std::copy(r.senC[k].begin(), r.senC[k].end(), senC + 3 * k);
    }
}

// Prepares (validation, gauge, active problem, groups) without solving. Returns 1 on success.
int oracle_prepare(oracle_ctx* ctx, const pb_solve_options* opts, pb_solve_report* report)
{
    const bool ok = prepare(ctx->o, opts);
    if (report) *report = ctx->o.report;
    return ok ? 1 : 0;
}

// Reduced system at the initial state (after oracle_prepare), dense 13-padded.
// matrix may be NULL (only rhs/cost). Returns primary block count, or -1 on failure.
int oracle_reduced_system(oracle_ctx* ctx, double damping, double* matrix, double* rhs, double* cost,
                          int32_t* singular)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return -1;
    try
    {
        Reduced R = buildReduced(o, o.init, damping);
        if (matrix)
        {
            std::vector<double> A;
            denseReduced(R, &A);
            std::copy(A.begin(), A.end(), matrix);
        }
        if (rhs)
            for (std::size_t i = 0; i < R.gradient.size(); ++i) rhs[i] = -R.gradient[i];
        if (cost) *cost = R.cost;
        if (singular) *singular = R.singular;
        return R.P;
    }
    catch (const std::exception&)
    {
        return -1;
    }
}

// One Gauss-Newton/LM step at the initial state: primary [P*13], eliminated [Ne*3], directional decrease.
int oracle_compute_step(oracle_ctx* ctx, double damping, double* primary, double* elim, double* dd,
                        double* rms_global, double* rms_point)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return -1;
    try
    {
        Reduced R = buildReduced(o, o.init, damping);
        if (R.singular) return -2;
        std::vector<double> p, e;
        double rel = 0.0;
        if (!solveReduced(R, o.act.cameraBlockCount, &p, &rel)) return -3;
        double dec = 0.0;
        if (!backSubstitute(o, p, R, &e, &dec)) return -4;
        if (primary) std::copy(p.begin(), p.end(), primary);
        if (elim) std::copy(e.begin(), e.end(), elim);
        if (dd) *dd = dec;
        const StepRms r = stepRms(o, o.init, p, e);
        if (rms_global) *rms_global = r.global;
        if (rms_point) *rms_point = r.point;
        return 0;
    }
    catch (const std::exception&)
    {
        return -5;
    }
}

// Analytic decrease of the objective at the initial state along an arbitrary step (primary [P*13], elim
// [Ne*3]): primary . b_c + sum_p elim_p . b_p, the form the back-substitution uses for the solved step
// (ReferenceSchur.cpp:1004-1007,1060-1062). Returns NaN on failure.
double oracle_decrease(oracle_ctx* ctx, const double* primary, const double* elim)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return std::numeric_limits<double>::quiet_NaN();
    try
    {
        Reduced R = buildReduced(o, o.init, 0.0);
        double dec = 0.0;
        for (std::size_t i = 0; i < R.directRhs.size(); ++i) dec += primary[i] * R.directRhs[i];
        for (std::size_t t = 0; t < o.pb.tracks.size(); ++t)
        {
            const int b = o.act.trackBlock[t];
            if (!o.act.activeTrack[t] || b < 0) continue;
            for (int r = 0; r < kE; ++r) dec += o.backSub[t].rhs[r] * elim[static_cast<std::size_t>(b) * kE + r];
        }
        return dec;
    }
    catch (const std::exception&)
    {
        return std::numeric_limits<double>::quiet_NaN();
    }
}

// Objective at the initial state moved by alpha * (primary, elim) through the retraction.
double oracle_cost_after_step(oracle_ctx* ctx, const double* primary, const double* elim, double alpha)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return std::numeric_limits<double>::quiet_NaN();
    State st = o.init;
    std::vector<double> p(primary, primary + static_cast<std::size_t>(o.act.primaryBlockCount) * kP);
    std::vector<double> e(elim, elim + static_cast<std::size_t>(o.act.trackBlockCount) * kE);
    applyStep(o, p, e, &st, alpha);
    return evaluateObjective(o, st);
}

// Intrinsic groups after preparation: per group 13 values each of params, prior, inverse sigma (or transition
// weight), lower, upper, and the enabled mask. Returns the group count (-1 when not prepared).
int oracle_groups(oracle_ctx* ctx, double* params, double* prior, double* inv_sigma, double* lower, double* upper,
                  uint32_t* enabled)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return -1;
    const auto& gs = o.init.groups;
    for (std::size_t g = 0; g < gs.size(); ++g)
    {
        uint32_t m = 0;
        for (int i = 0; i < kP; ++i)
        {
            if (params) params[g * kP + i] = gs[g].params[i];
            if (prior) prior[g * kP + i] = gs[g].prior[i];
            if (inv_sigma) inv_sigma[g * kP + i] = gs[g].invSigma[i];
            if (lower) lower[g * kP + i] = gs[g].lower[i];
            if (upper) upper[g * kP + i] = gs[g].upper[i];
            if (gs[g].enabled[i]) m |= 1u << i;
        }
        if (enabled) enabled[g] = m;
    }
    return static_cast<int>(gs.size());
}

double oracle_initial_cost(oracle_ctx* ctx)
{
    Oracle& o = ctx->o;
    if (!o.prepared) return std::numeric_limits<double>::quiet_NaN();
    return evaluateObjective(o, o.init);
}

// --- per-observation primitives for Jacobian and loss checks -------------------------------------
int oracle_project(const double* R, const double* C, const double* slots, int conv, const double* X, double* uv,
                   double* Jcam, double* Jslot)
{
    Camera cam;
    std::copy(R, R + 9, cam.R.begin());
    std::copy(C, C + 3, cam.C.begin());
    std::copy(slots, slots + kP, cam.p.begin());
    cam.conv = conv;
    ProjectionOut pr;
    const Vec3 x{X[0], X[1], X[2]};
    if (!project(cam, x, true, &pr)) return 0;
    uv[0] = pr.uv[0];
    uv[1] = pr.uv[1];
    if (Jcam) std::copy(pr.Jcam, pr.Jcam + 6, Jcam);
    if (Jslot) std::copy(pr.Jslot, pr.Jslot + 26, Jslot);
    return 1;
}

int oracle_robust_loss(double s, int loss, double scale, double* cost, double* weight)
{
    return robustLoss(s, loss, scale, cost, weight) ? 1 : 0;
}

double oracle_point_scale(const double* X) { return pointScale({X[0], X[1], X[2]}); }

// Reference gauge tangent basis of a scale camera at world centre `scale` around the anchor at `anchor`
// (referenceGaugeTargetTangentBasis, Assembly.cpp:69-99), written as world vectors: basis[0..2] the first,
// basis[3..5] the second (the scale camera's Jacobian columns 3 and 4 are Jp times these, Assembly.cpp:295-306).
void oracle_gauge_basis(const double* anchor, const double* scale, double* basis)
{
    Options opt;
    opt.gaugeAnchor = 0;
    opt.gaugeScale = 1;
    opt.gaugeBaseline = 1.0;
    State st;
    st.cams.resize(2);
    st.cams[0].C = {anchor[0], anchor[1], anchor[2]};
    st.cams[1].C = {scale[0], scale[1], scale[2]};
    const auto b = gaugeBasis(opt, st);
    for (int k = 0; k < 2; ++k)
    {
        const Vec3 w = targetCoords(b[k]);
        for (int i = 0; i < 3; ++i) basis[3 * k + i] = w[i];
    }
}

// New world centre of the scale camera after the tangent step (first, second) at the given baseline
// (applyReferenceGaugeTangentStep, Assembly.cpp:101-128).
void oracle_gauge_step(const double* anchor, const double* scale, double baseline, double first, double second,
                       double* center)
{
    Options opt;
    opt.gaugeAnchor = 0;
    opt.gaugeScale = 1;
    opt.gaugeBaseline = baseline;
    State st;
    st.cams.resize(2);
    st.cams[0].C = {anchor[0], anchor[1], anchor[2]};
    st.cams[1].C = {scale[0], scale[1], scale[2]};
    const Vec3 c = gaugeTangentStep(opt, st, first, second);
    for (int i = 0; i < 3; ++i) center[i] = c[i];
}

void oracle_retract(double* R, double* C, const double* delta, int reference_param)
{
    Mat3 r;
    Vec3 c;
    std::copy(R, R + 9, r.begin());
    std::copy(C, C + 3, c.begin());
    if (reference_param)
        retractReference(&r, &c, delta);
    else
        retractExp(&r, &c, delta);
    std::copy(r.begin(), r.end(), R);
    std::copy(c.begin(), c.end(), C);
}

// Linearisation of one observation: residual[2], Jp[6], Jpose[12], Jint[26], weight, cost.
int oracle_linearize(const double* R, const double* C, const double* slots, int conv, const double* X, double u,
                     double v, double weight, double scale, int loss, double loss_scale, int reference_param,
                     uint32_t active_mask, double* residual, double* Jp, double* Jpose, double* Jint,
                     double* normal_weight, double* robust_cost)
{
    Camera cam;
    std::copy(R, R + 9, cam.R.begin());
    std::copy(C, C + 3, cam.C.begin());
    std::copy(slots, slots + kP, cam.p.begin());
    cam.conv = conv;
    Observation o;
    o.cam = 0;
    o.u = u;
    o.v = v;
    o.weight = weight;
    o.scale = scale;
    SlotMask act{};
    for (int i = 0; i < kP; ++i) act[i] = (active_mask >> i) & 1u;
    Linearization L;
    if (!linearize(cam, {X[0], X[1], X[2]}, o, loss, loss_scale, reference_param != 0, act, &L)) return 0;
    std::copy(L.r, L.r + 2, residual);
    std::copy(L.Jp, L.Jp + 6, Jp);
    std::copy(L.Jpose, L.Jpose + 12, Jpose);
    std::copy(L.Jint, L.Jint + 26, Jint);
    *normal_weight = L.w;
    *robust_cost = L.cost;
    return 1;
}


void oracle_default_quality_options(pb_quality_options* q)
{
    q->enable_point_filter = 1;
    q->filter_max_reproj_error = 2.5;
    q->filter_sigma_factor = 3.0;
    q->enable_quality_gate = 1;
    q->max_accepted_rms_growth = 1.25;
    q->min_accepted_valid_track_ratio = 0.60;
}

// finalizeBundleAdjustResult (BundleAdjustQuality.cpp:394-502) + the image checks of resultFailsQualityGate
// (solver.cpp:123-171) on the oracle's last solve.
int oracle_quality(oracle_ctx* ctx, const pb_quality_options* q, pb_quality_report* rep, double* rmsBefore,
                   double* rmsAfter, int32_t* validOut)
{
    Oracle& o = ctx->o;
    const double inf = std::numeric_limits<double>::infinity();
    const std::size_t nT = o.pb.tracks.size();
    std::memset(rep, 0, sizeof(*rep));
    const State& st = o.state.cams.empty() ? o.init : o.state;
    const std::vector<Camera> refined = (o.prepared && !st.cams.empty()) ? effectiveCameras(o.pb, o.act, st)
                                                                         : o.pb.cams;
    std::vector<double> before(nT), after(nT, inf), candidates;
    std::vector<char> valid(nT, 0);
    for (std::size_t t = 0; t < nT; ++t)
    {
        const Track& tr = o.pb.tracks[t];
        before[t] = strictTrackRms(o.pb.cams, tr, tr.X);
        if (!o.prepared || !o.act.activeTrack[t]) continue;
        const Vec3& X = st.points.empty() ? tr.X : st.points[t];
        after[t] = strictTrackRms(refined, tr, X);
        valid[t] = std::isfinite(after[t]) ? 1 : 0;
        if (valid[t]) candidates.push_back(after[t]);
    }
    const double threshold = q->enable_point_filter
                                 ? adaptiveThreshold(candidates, q->filter_max_reproj_error, q->filter_sigma_factor)
                                 : inf;
    int optimized = 0;
    double sumBefore = 0.0, sumAfter = 0.0;
    for (std::size_t t = 0; t < nT; ++t)
    {
        if (valid[t] && after[t] > threshold) valid[t] = 0;
        if (!valid[t]) continue;
        ++optimized;
        sumBefore += before[t];
        sumAfter += after[t];
    }
    rep->total_tracks = static_cast<int32_t>(nT);
    rep->optimized_tracks = optimized;
    rep->mean_rms_before = optimized > 0 ? sumBefore / optimized : inf;
    rep->mean_rms_after = optimized > 0 ? sumAfter / optimized : inf;
    rep->valid_track_ratio = nT > 0 ? double(optimized) / double(nT) : 0.0;
    rep->median_rms_after = medianOr(candidates, 0.0);
    rep->filter_threshold = threshold;
    rep->solve_status = o.report.solve_status;
    rep->solution_usable = o.report.solution_usable;
    if (optimized == 0 &&
        (rep->solve_status == PB_SOLVE_SUCCESS || rep->solve_status == PB_SOLVE_NO_CONVERGENCE))
    {
        rep->solve_status = PB_SOLVE_NUMERICAL_FAILURE;
        rep->solution_usable = 0;
        std::snprintf(rep->message, sizeof(rep->message), "final quality check kept no valid track");
    }
    if (q->enable_quality_gate)
    {
        const char* why = nullptr;
        const double growth = std::max(0.0, q->max_accepted_rms_growth);
        if (!rep->solution_usable)
            why = "quality gate rejected a non-usable solution";
        else if (nT > 0 && !std::isfinite(rep->mean_rms_after))
            why = "quality gate rejected a non-finite final RMS";
        else if (nT > 0 && rep->valid_track_ratio < std::max(0.0, q->min_accepted_valid_track_ratio))
            why = "quality gate rejected a low valid-track ratio";
        else if (nT > 0 && growth > 0.0 && std::isfinite(rep->mean_rms_before))
        {
            if (rep->mean_rms_before > 1e-12 && rep->mean_rms_after > rep->mean_rms_before * growth)
                why = "quality gate rejected excessive reprojection RMS growth";
            else if (rep->mean_rms_before <= 1e-12 && rep->mean_rms_after > 1e-9)
                why = "quality gate rejected growth from a zero-residual input";
        }
        if (why)
        {
            rep->quality_gate_rejected = 1;
            std::snprintf(rep->message, sizeof(rep->message), "%s", why);
        }
    }
    for (std::size_t t = 0; t < nT; ++t)
    {
        if (rmsBefore) rmsBefore[t] = before[t];
        if (rmsAfter) rmsAfter[t] = after[t];
        if (validOut) validOut[t] = valid[t];
    }
    return rep->quality_gate_rejected;
}

}  // extern "C"
