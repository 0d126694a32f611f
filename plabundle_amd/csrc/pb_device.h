// pb_device.h - device-side frame-camera model and per-observation linearisation for gfx950.
//
// One observation of a world point X by a frame camera (R camera-to-world, centre C) gives the
// residual r = pi(X) - (u, v), the point Jacobian Jp (2x3), the pose Jacobian (2x6: rotation, world
// centre) and the Jacobian of the group's active intrinsic slots. Two parameterisations
// (Assembly.cpp:24-29): the reference one of the online Schur without survey constraints (rotation in
// the D = diag(1,-1,-1) target frame, Euler retraction, point step scaled by s/atan s) and the default
// one (world-frame left rotation increment, Exp retraction, unscaled point step). The math follows
// BundleAdjustPlaMatrixProjection.cpp:176-285 (residual, Jacobian layout, weights) and
// CameraState.cpp:387-481 (13-slot intrinsic layout and the focal/aspect chain rule).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define PB_HD __host__ __device__ __forceinline__

namespace pbdev
{
    constexpr int kSlots = 13;
    constexpr int kMaxActive = 13;  // active intrinsic slots per group handled on device (all 13: 3 units of 6)

    // Effective camera used by every kernel of one pass (pose + intrinsics after the shared-group overlay).
    struct Camera
    {
        double R[9];
        double C[3];
        double fx, fy, cx, cy;
        double k1, k2, k3, k4, p1, p2, p3, p4, b2;
        double basis[6];  // world-frame tangent directions of the gauge scale camera (D * basis)
        int32_t metashape;
        int32_t tangent;  // 1 for the gauge scale camera (centre moves on the baseline sphere)
        int32_t pad[2];
    };

    struct Linearization
    {
        double r[2];
        double Jp[6];    // 2x3, row-major
        double Jc[12];   // 2x6, row-major [rotation | centre]
        double Ji[2 * kMaxActive];  // 2 x K compact intrinsic Jacobian, row-major with stride kMaxActive
        double w;        // normal weight (whitened weight * robust weight)
        double cost;     // robust cost
    };

    PB_HD bool finite(double v) { return __builtin_isfinite(v); }

    // s / atan(s) with s = |X| / 1000 (BundleAdjustPlaMatrixProjection.cpp:15-24).
    PB_HD double pointScale(double x, double y, double z)
    {
        const double s = sqrt(x * x + y * y + z * z) / 1000.0;
        if (!finite(s) || s == 0.0) return 1.0;
        return s / atan(s);
    }

    // Robust loss (BundleAdjustPlaMatrixProjection.cpp:75-109): returns cost and IRLS weight.
    PB_HD bool robust(double s, int loss, double scale, double& cost, double& weight)
    {
        if (!finite(s) || s < 0.0) return false;
        const double d2 = scale * scale;
        if (loss == 1)
        {
            if (s <= d2)
            {
                cost = 0.5 * s;
                weight = 1.0;
            }
            else
            {
                const double n = sqrt(s);
                cost = scale * n - 0.5 * d2;
                weight = scale / n;
            }
        }
        else if (loss == 2)
        {
            const double q = s / d2;
            cost = 0.5 * d2 * log1p(q);
            weight = 1.0 / (1.0 + q);
        }
        else
        {
            cost = 0.5 * s;
            weight = 1.0;
        }
        return true;
    }

    // Distortion on the normalised plane: (x, y) -> (xd, yd) and, optionally, d(xd,yd)/d(x,y).
    struct Distorted
    {
        double xd, yd;
        double r2, rad, tx0, ty0, T;
        double dxx, dxy, dyx, dyy;
    };

    template <bool kJacobian>
    PB_HD Distorted distort(const Camera& c, double x, double y)
    {
        Distorted d;
        const double r2 = x * x + y * y;
        d.r2 = r2;
        d.rad = 1.0 + r2 * (c.k1 + r2 * (c.k2 + r2 * (c.k3 + r2 * c.k4)));
        if (c.metashape)
        {
            d.tx0 = c.p1 * (r2 + 2.0 * x * x) + 2.0 * c.p2 * x * y;
            d.ty0 = c.p2 * (r2 + 2.0 * y * y) + 2.0 * c.p1 * x * y;
            d.T = 1.0 + c.p3 * r2 + c.p4 * r2 * r2;
        }
        else
        {
            d.tx0 = 2.0 * c.p1 * x * y + c.p2 * (r2 + 2.0 * x * x);
            d.ty0 = c.p1 * (r2 + 2.0 * y * y) + 2.0 * c.p2 * x * y;
            d.T = 1.0;
        }
        d.xd = x * d.rad + d.tx0 * d.T;
        d.yd = y * d.rad + d.ty0 * d.T;
        if (kJacobian)
        {
            const double drad = c.k1 + r2 * (2.0 * c.k2 + r2 * (3.0 * c.k3 + r2 * 4.0 * c.k4));
            double txx, txy, tyx, tyy;
            if (c.metashape)
            {
                const double dT = 2.0 * (c.p3 + 2.0 * c.p4 * r2);
                txx = (6.0 * c.p1 * x + 2.0 * c.p2 * y) * d.T + d.tx0 * dT * x;
                txy = (2.0 * c.p1 * y + 2.0 * c.p2 * x) * d.T + d.tx0 * dT * y;
                tyx = (2.0 * c.p2 * x + 2.0 * c.p1 * y) * d.T + d.ty0 * dT * x;
                tyy = (6.0 * c.p2 * y + 2.0 * c.p1 * x) * d.T + d.ty0 * dT * y;
            }
            else
            {
                txx = 2.0 * c.p1 * y + 6.0 * c.p2 * x;
                txy = 2.0 * c.p1 * x + 2.0 * c.p2 * y;
                tyx = txy;
                tyy = 6.0 * c.p1 * y + 2.0 * c.p2 * x;
            }
            d.dxx = d.rad + 2.0 * x * x * drad + txx;
            d.dxy = 2.0 * x * y * drad + txy;
            d.dyx = 2.0 * x * y * drad + tyx;
            d.dyy = d.rad + 2.0 * y * y * drad + tyy;
        }
        return d;
    }

    // Projection only (cost passes). Forward depth must exceed 1e-9.
    PB_HD bool project(const Camera& c, double X0, double X1, double X2, double& u, double& v)
    {
        const double d0 = X0 - c.C[0], d1 = X1 - c.C[1], d2 = X2 - c.C[2];
        const double zc = c.R[2] * d0 + c.R[5] * d1 + c.R[8] * d2;
        if (!(zc > 1e-9)) return false;
        const double iz = 1.0 / zc;
        const double x = (c.R[0] * d0 + c.R[3] * d1 + c.R[6] * d2) * iz;
        const double y = (c.R[1] * d0 + c.R[4] * d1 + c.R[7] * d2) * iz;
        const Distorted d = distort<false>(c, x, y);
        u = c.fx * d.xd + c.b2 * d.yd + c.cx;
        v = c.fy * d.yd + c.cy;
        return finite(u) && finite(v);
    }

    // Derivative of (u, v) with respect to one intrinsic slot.
    PB_HD void slotDerivative(const Camera& c, const Distorted& d, double x, double y, int slot,
                                                   double& du, double& dv)
    {
        double ex = 0.0, ey = 0.0;
        switch (slot)
        {
        case 0:  // fx at fixed aspect
            du = d.xd;
            dv = (c.fy / c.fx) * d.yd;
            return;
        case 1:  // log(fy / fx)
            du = 0.0;
            dv = c.fy * d.yd;
            return;
        case 2:
            du = 1.0;
            dv = 0.0;
            return;
        case 3:
            du = 0.0;
            dv = 1.0;
            return;
        case 9:  // skew b2
            du = d.yd;
            dv = 0.0;
            return;
        case 4: ex = x * d.r2; ey = y * d.r2; break;
        case 5: ex = x * d.r2 * d.r2; ey = y * d.r2 * d.r2; break;
        case 6: ex = x * d.r2 * d.r2 * d.r2; ey = y * d.r2 * d.r2 * d.r2; break;
        case 10:
        {
            const double r8 = (d.r2 * d.r2) * (d.r2 * d.r2);
            ex = x * r8;
            ey = y * r8;
            break;
        }
        case 7:
            if (c.metashape) { ex = (d.r2 + 2.0 * x * x) * d.T; ey = 2.0 * x * y * d.T; }
            else { ex = 2.0 * x * y; ey = d.r2 + 2.0 * y * y; }
            break;
        case 8:
            if (c.metashape) { ex = 2.0 * x * y * d.T; ey = (d.r2 + 2.0 * y * y) * d.T; }
            else { ex = d.r2 + 2.0 * x * x; ey = 2.0 * x * y; }
            break;
        case 11:
            if (c.metashape) { ex = d.tx0 * d.r2; ey = d.ty0 * d.r2; }
            break;
        case 12:
            if (c.metashape) { ex = d.tx0 * d.r2 * d.r2; ey = d.ty0 * d.r2 * d.r2; }
            break;
        default: break;
        }
        du = c.fx * ex + c.b2 * ey;
        dv = c.fy * ey;
    }

    // Full linearisation of one observation. nActive/activeSlots select the intrinsic columns.
    // KM: compile-time bound on the active intrinsic slots (0, 6, 12 or 13) so that Ji lives in registers.
    template <int KM>
    PB_HD bool linearize(const Camera& c, double X0, double X1, double X2, double ou, double ov,
                                              double ow, int loss, double lscale, int nActive,
                                              const int8_t* activeSlots, Linearization& L, bool refParam = true)
    {
        const double d0 = X0 - c.C[0], d1 = X1 - c.C[1], d2 = X2 - c.C[2];
        const double zc = c.R[2] * d0 + c.R[5] * d1 + c.R[8] * d2;
        if (!(zc > 1e-9)) return false;
        const double iz = 1.0 / zc;
        const double x = (c.R[0] * d0 + c.R[3] * d1 + c.R[6] * d2) * iz;
        const double y = (c.R[1] * d0 + c.R[4] * d1 + c.R[7] * d2) * iz;
        const Distorted d = distort<true>(c, x, y);
        const double u = c.fx * d.xd + c.b2 * d.yd + c.cx;
        const double v = c.fy * d.yd + c.cy;
        L.r[0] = u - ou;
        L.r[1] = v - ov;
        // d(u,v)/d(x,y)
        const double ux = c.fx * d.dxx + c.b2 * d.dyx;
        const double uy = c.fx * d.dxy + c.b2 * d.dyy;
        const double vx = c.fy * d.dyx;
        const double vy = c.fy * d.dyy;
        // camera-frame Jacobian rows, then world point Jacobian Jp = Jcam * R^T
        const double cu[3] = {ux * iz, uy * iz, -(ux * x + uy * y) * iz};
        const double cv[3] = {vx * iz, vy * iz, -(vx * x + vy * y) * iz};
        double jp[6];
#pragma unroll
        for (int w = 0; w < 3; ++w)
        {
            jp[w] = cu[0] * c.R[w * 3 + 0] + cu[1] * c.R[w * 3 + 1] + cu[2] * c.R[w * 3 + 2];
            jp[3 + w] = cv[0] * c.R[w * 3 + 0] + cv[1] * c.R[w * 3 + 1] + cv[2] * c.R[w * 3 + 2];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
        {
            const double* p = jp + 3 * a;
            double* J = L.Jc + 6 * a;
            // reference: rotation in the target frame, J_rot = -Jp [d]x D; default: world-frame left
            // increment, J_rot = Jp [d]x (the two differ in the sign of the first column)
            J[0] = refParam ? -p[1] * d2 + p[2] * d1 : p[1] * d2 - p[2] * d1;
            J[1] = -p[0] * d2 + p[2] * d0;
            J[2] = p[0] * d1 - p[1] * d0;
            if (c.tangent)
            {
                J[3] = -(p[0] * c.basis[0] + p[1] * c.basis[1] + p[2] * c.basis[2]);
                J[4] = -(p[0] * c.basis[3] + p[1] * c.basis[4] + p[2] * c.basis[5]);
                J[5] = 0.0;
            }
            else
            {
                J[3] = -p[0];
                J[4] = -p[1];
                J[5] = -p[2];
            }
        }
        const double ps = refParam ? pointScale(X0, X1, X2) : 1.0;
#pragma unroll
        for (int i = 0; i < 6; ++i) L.Jp[i] = jp[i] * ps;
        // fixed trip count so that Ji stays in registers (a runtime bound would index it dynamically)
#pragma unroll
        for (int k = 0; k < KM; ++k)
        {
            double du = 0.0, dv = 0.0;
            if (k < nActive) slotDerivative(c, d, x, y, activeSlots[k], du, dv);
            L.Ji[k] = du;
            L.Ji[kMaxActive + k] = dv;
        }
        const double s = ow * (L.r[0] * L.r[0] + L.r[1] * L.r[1]);
        double cost, rw;
        if (!robust(s, loss, lscale, cost, rw)) return false;
        L.w = ow * rw;
        L.cost = cost;
        return finite(L.w) && finite(cost) && finite(u) && finite(v);
    }

    // Second derivative rho''(s) of the per-observation cost of robust() (cost = rho(s), weight = 2 rho'(s)).
    PB_HD double robustCurvature(double s, int loss, double scale)
    {
        const double d2 = scale * scale;
        if (loss == 1) return s <= d2 ? 0.0 : -0.25 * scale / (s * sqrt(s));
        if (loss == 2)
        {
            const double q = 1.0 + s / d2;
            return -0.5 / (d2 * q * q);
        }
        return 0.0;
    }

    // Symmetric square root Q (q00, q01, q11) of an observation's exact weight matrix for the adjoint:
    // the gradient of the cost with respect to the residual is w_n r (w_n = ow * 2 rho'), and
    // M = d(w_n r)/dr = w_n I + 4 ow^2 rho''(s) r r^T; its eigenvalue along r is clamped at zero so that the
    // adjoint system stays positive semi-definite (it is never negative for least squares and Huber).
    PB_HD void adjointWeightRoot(const Linearization& L, double ow, int loss, double scale, double q[3])
    {
        const double n2 = L.r[0] * L.r[0] + L.r[1] * L.r[1];
        const double sw = sqrt(fmax(L.w, 0.0));
        if (!(n2 > 0.0))
        {
            q[0] = sw;
            q[1] = 0.0;
            q[2] = sw;
            return;
        }
        const double c = 4.0 * ow * ow * robustCurvature(ow * n2, loss, scale);
        const double sm = sqrt(fmax(0.0, L.w + c * n2));
        const double a = L.r[0] * L.r[0] / n2, b = L.r[0] * L.r[1] / n2, d = L.r[1] * L.r[1] / n2;
        q[0] = sw * (1.0 - a) + sm * a;
        q[1] = (sm - sw) * b;
        q[2] = sw * (1.0 - d) + sm * d;
    }

    // Symmetric root of a survey constraint's exact weight matrix d(w r)/dr = w I + 4 rho''(s) r r^T (Huber on
    // s = |r|^2, 1 or 3 rows): Q = a I + (b - a) r r^T / s with a = sqrt(w), b = sqrt(max(0, w + 4 rho'' s))
    // (zero along r on the linear branch). Returns the pair (a, b - a) scaled into q[0], q[1].
    PB_HD void constraintWeightRoot(double s, double w, double delta, double q[2])
    {
        const double a = sqrt(fmax(w, 0.0));
        const double b = sqrt(fmax(0.0, w + 4.0 * robustCurvature(s, 1, delta) * s));
        q[0] = a;
        q[1] = s > 0.0 ? (b - a) / s : 0.0;
    }

    // J <- Q J for every Jacobian row pair, weight 1: the normal equations of the transformed rows are
    // sum J^T M J.
    template <int KM>
    PB_HD void applyWeightRoot(const double q[3], Linearization& L)
    {
        auto tr = [&](double& j0, double& j1) {
            const double a = j0, b = j1;
            j0 = q[0] * a + q[1] * b;
            j1 = q[1] * a + q[2] * b;
        };
#pragma unroll
        for (int i = 0; i < 3; ++i) tr(L.Jp[i], L.Jp[3 + i]);
#pragma unroll
        for (int i = 0; i < 6; ++i) tr(L.Jc[i], L.Jc[6 + i]);
#pragma unroll
        for (int k = 0; k < KM; ++k) tr(L.Ji[k], L.Ji[kMaxActive + k]);
        L.w = 1.0;
    }

    // Reference pose retraction C2W' = D E^T D C2W, E = Rx(a) Ry(b) Rz(c) (Assembly.cpp:130-179).
    PB_HD void retractRotation(const double* R, double a, double b, double cz, double* out)
    {
        const double sx = sin(a), cx = cos(a), sy = sin(b), cy = cos(b), sz = sin(cz), ccz = cos(cz);
        // E = Rx Ry Rz
        const double E[9] = {cy * ccz, -cy * sz, sy,
                             sx * sy * ccz + cx * sz, -sx * sy * sz + cx * ccz, -sx * cy,
                             -cx * sy * ccz + sx * sz, cx * sy * sz + sx * ccz, cx * cy};
        const double sg[3] = {1.0, -1.0, -1.0};
        double M[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) M[r * 3 + k] = sg[r] * E[k * 3 + r] * sg[k];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                out[r * 3 + k] = M[r * 3 + 0] * R[0 * 3 + k] + M[r * 3 + 1] * R[1 * 3 + k] + M[r * 3 + 2] * R[2 * 3 + k];
    }

    // Default pose retraction (PlaCamera applyPoseDelta, CameraState.cpp:374-380): C2W' = Exp([w]x) C2W
    // with the Rodrigues form of Exp.
    PB_HD void retractExp(const double* R, double wx, double wy, double wz, double* out)
    {
        const double t2 = wx * wx + wy * wy + wz * wz;
        double E[9];
        if (t2 < 1e-20)
        {
            E[0] = 1.0; E[1] = -wz; E[2] = wy;
            E[3] = wz; E[4] = 1.0; E[5] = -wx;
            E[6] = -wy; E[7] = wx; E[8] = 1.0;
        }
        else
        {
            const double t = sqrt(t2);
            const double sc = sin(t) / t, cc = (1.0 - cos(t)) / t2;
            E[0] = 1.0 - cc * (wy * wy + wz * wz); E[1] = cc * wx * wy - sc * wz; E[2] = cc * wx * wz + sc * wy;
            E[3] = cc * wx * wy + sc * wz; E[4] = 1.0 - cc * (wx * wx + wz * wz); E[5] = cc * wy * wz - sc * wx;
            E[6] = cc * wx * wz - sc * wy; E[7] = cc * wy * wz + sc * wx; E[8] = 1.0 - cc * (wx * wx + wy * wy);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                out[r * 3 + k] = E[r * 3 + 0] * R[0 * 3 + k] + E[r * 3 + 1] * R[1 * 3 + k] + E[r * 3 + 2] * R[2 * 3 + k];
    }

    // Rig camera pose (composeRigCameras, as runConcreteBackend applies it, solver.cpp:282): camera-to-world
    // R = R_rig R_sensor, centre C = C_rig + R_rig c_sensor.
    PB_HD void composeRig(const double* capR, const double* capC, const double* senR, const double* senC, double* R,
                          double* C)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                R[i * 3 + j] = capR[i * 3 + 0] * senR[0 * 3 + j] + capR[i * 3 + 1] * senR[1 * 3 + j] +
                               capR[i * 3 + 2] * senR[2 * 3 + j];
            C[i] = capC[i] + (capR[i * 3 + 0] * senC[0] + capR[i * 3 + 1] * senC[1] + capR[i * 3 + 2] * senC[2]);
        }
    }

    // Chain maps of a rig camera (cameraPosePrimaryTerms, Assembly.cpp:224-281): with the camera's pose Jacobian
    // J = [rotation | centre] in world-frame increments, the capture's Jacobian is J Mk and the sensor's J Ms
    // (6 x 6, row-major). Mk = [[I, 0], [K, I]] with column a of K the centre's motion e_a x l under a rotation of
    // the rig about axis a (lever arm l = R_rig c_sensor); Ms = diag(R_rig, R_rig) (sensor increments in rig axes).
    PB_HD void rigMaps(const double* capR, const double* senC, double* Mk, double* Ms)
    {
        double l[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) l[i] = capR[i * 3 + 0] * senC[0] + capR[i * 3 + 1] * senC[1] + capR[i * 3 + 2] * senC[2];
#pragma unroll
        for (int i = 0; i < 36; ++i)
        {
            Mk[i] = 0.0;
            Ms[i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) Mk[i * 6 + i] = 1.0;
        // e_a x l for a = 0, 1, 2 as the columns of K (rows 3..5, columns 0..2)
        Mk[4 * 6 + 0] = -l[2];
        Mk[5 * 6 + 0] = l[1];
        Mk[3 * 6 + 1] = l[2];
        Mk[5 * 6 + 1] = -l[0];
        Mk[3 * 6 + 2] = -l[1];
        Mk[4 * 6 + 2] = l[0];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                Ms[i * 6 + j] = capR[i * 3 + j];
                Ms[(3 + i) * 6 + 3 + j] = capR[i * 3 + j];
            }
    }

    // Huber loss of a survey constraint (BundleAdjustPlaMatrixConstraints.cpp:17-29): s = |r|^2.
    PB_HD bool huber(double s, double delta, double& cost, double& weight)
    {
        return finite(delta) && delta > 0.0 && robust(s, 1, delta, cost, weight) && finite(cost) && finite(weight);
    }

    // Symmetric 3x3 inverse (row-major) with the singularity test |det| <= tol max|m|^3 (tol = 1e-15).
    PB_HD bool inverse3(const double* m, double* inv, double tol = 1e-15)
    {
        double mx = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i)
        {
            if (!finite(m[i])) return false;
            mx = fmax(mx, fabs(m[i]));
        }
        const double c00 = m[4] * m[8] - m[5] * m[7];
        const double c01 = m[5] * m[6] - m[3] * m[8];
        const double c02 = m[3] * m[7] - m[4] * m[6];
        const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
        if (!finite(det) || !(fabs(det) > tol * mx * mx * mx)) return false;
        const double id = 1.0 / det;
        inv[0] = c00 * id;
        inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
        inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
        inv[3] = c01 * id;
        inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
        inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
        inv[6] = c02 * id;
        inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
        inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            if (!finite(inv[i])) return false;
        return true;
    }
}  // namespace pbdev
