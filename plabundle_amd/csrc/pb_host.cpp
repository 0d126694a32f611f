// pb_host.cpp - host side of the MI355X frame-BA backend: the C ABI of include/pb_hip.h.
//
// Host work per solve (C++20):
//   1. options validation and gauge normalisation .... BundleAdjustValidation.cpp:371-896 (image-only subset)
//   2. active problem and block numbering ............ BundleAdjustPlaMatrixProblem.cpp:29-278
//   3. shared-intrinsic groups ....................... BundleAdjustPlaMatrixModel.cpp:84-231
//   4. device layout: track-sorted observations, camera-sorted permutation, cross-block entries,
//      Schur pair runs and the block-sparse reduced-system pattern (built once per problem)
//   5. the online-Schur Gauss-Newton/Armijo controller BundleAdjustPlaMatrix.cpp:162-631, issuing the
//      kernels of pb_kernels.hip on one HIP stream and reading back a few scalars per stage.
#include "../../include/pb_hip.h"
#include "pb_device.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    using Clock = std::chrono::steady_clock;
    constexpr double kInf = std::numeric_limits<double>::infinity();

    struct HipError : std::runtime_error
    {
        hipError_t code;
        HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
    };

    void check(hipError_t e, const char* where)
    {
        if (e != hipSuccess)
        {
            std::string msg = std::string(where) + ": " + hipGetErrorString(e);
            throw HipError(e, msg.c_str());
        }
    }

    double seconds(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

    // ------------------------------------------------------------------------------------------
    // Input copy of one problem.
    struct Input
    {
        int nCam = 0, nTracks = 0;
        std::vector<double> R, C, slots, ref;
        std::vector<int32_t> conv, group;
        std::vector<int32_t> imageSize;  // [nCam*2] width, height, or empty
        bool hasGroup = false;
        std::vector<double> X;
        std::vector<int64_t> off;
        std::vector<int32_t> obsCam;
        std::vector<double> uv, weight, scale;
        std::vector<int32_t> fixedCams, fixedTracks;
        int gaugePolicy = PB_GAUGE_AUTO_ANCHOR, gaugeAnchor = -1, gaugeScale = -1;
        double gaugeBaseline = 0.0;
        // survey constraints (per-track control points and laser planes), in caller order per kind
        struct Constraint
        {
            int track = -1;
            bool plane = false;
            int index = -1;     // position in the caller's list of its kind
            double A[9] = {};   // control point: whitening matrix (row-major); laser plane: normal in row 0
            double target[3] = {};
            double weight = 1.0;
        };
        std::vector<Constraint> cons;
        int nControl = 0, nPlanes = 0;
        // camera rig: capture poses (rig-to-world), sensor poses (camera-to-rig), fixed flags and each camera's
        // binding; the bound cameras' R and C above hold the composed poses
        int nCap = 0, nSen = 0;
        std::vector<double> capR, capC, senR, senC;
        std::vector<int32_t> capFixed, senFixed, camCapture, camSensor;
        bool rig() const { return nCap > 0 || nSen > 0; }
    };

    double sanitizedWeight(double w) { return std::isfinite(w) && w > 0.0 ? w : 0.0; }
    double sanitizedScale(double s) { return std::isfinite(s) && s > 0.0 ? s : 1.0; }

    bool usable(const Input& in, int64_t i)
    {
        const int c = in.obsCam[i];
        return c >= 0 && c < in.nCam && std::isfinite(in.uv[2 * i]) && std::isfinite(in.uv[2 * i + 1]) &&
               sanitizedWeight(in.weight[i]) > 0.0;
    }

    pbdev::Camera inputCamera(const Input& in, int c)
    {
        pbdev::Camera k{};
        std::copy_n(&in.R[9 * c], 9, k.R);
        std::copy_n(&in.C[3 * c], 3, k.C);
        const double* s = &in.slots[13 * c];
        k.fx = s[0];
        k.fy = s[0] * std::exp(s[1]);
        k.cx = s[2];
        k.cy = s[3];
        k.k1 = s[4];
        k.k2 = s[5];
        k.k3 = s[6];
        k.p1 = s[7];
        k.p2 = s[8];
        k.b2 = s[9];
        k.k4 = s[10];
        k.metashape = in.conv[c] == PB_BROWN_METASHAPE;
        k.p3 = k.metashape ? s[11] : 0.0;
        k.p4 = k.metashape ? s[12] : 0.0;
        return k;
    }

    uint32_t supportedMask(int conv) { return conv == PB_BROWN_METASHAPE ? 0x1FFFu : 0x7FFu; }

    // ------------------------------------------------------------------------------------------
    // Host preparation of one solve.
    struct Prepared
    {
        std::vector<int32_t> fixedCams;
        std::vector<int32_t> fixedTracks;  // sorted
        int anchor = -1, scaleCam = -1;
        double baseline = 0.0;
        // active problem (bit-exact with prepareActiveProblem)
        std::vector<char> activeTrack;
        std::vector<int32_t> cameraBlock, intrinsicBlock, groupOfCam, trackBlock;
        int cameraBlocks = 0, intrinsicBlocks = 0, primaryBlocks = 0, trackBlocks = 0;
        int activeTracks = 0, observations = 0, rejectedInitial = 0;
        // groups
        int G = 0;
        std::vector<double> gParams, gPrior, gInvSigma, gLower, gUpper;
        std::vector<uint32_t> gEnabled;
        std::vector<int8_t> gActive;
        std::vector<int32_t> gK, gUnit, slotToCompact;
        // units
        int nUnits = 0, globalDim = 0;
        std::vector<int32_t> unitWidth;
        // camera rig, in the reference's numbering (BundleAdjustPlaMatrixProblem.cpp:157-199): the primary block of
        // each refined capture and sensor (-1 otherwise) and per camera its capture's and sensor's block. The
        // reference's camera blocks are then the rig blocks; cameraBlock / cameraBlocks above number the
        // device's camera-level pose units instead (cameras whose capture or sensor is refined).
        bool rig = false;
        int rigCapBlocks = 0, rigSenBlocks = 0;
        std::vector<int32_t> rigCapBlock, rigSenBlock, rigCapBlockByCam, rigSenBlockByCam;
        int refCameraBlocks() const { return rig ? rigCapBlocks + rigSenBlocks : cameraBlocks; }
    };

    // Two distinct cameras with a finite pixel and a positive weight (hasUsableTrackObservations,
    // BundleAdjustValidation.cpp:94-106; only the sign of the camera index is checked there).
    bool trackUsableForGauge(const Input& in, int t)
    {
        int first = -1;
        for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
        {
            if (in.obsCam[i] < 0 || !std::isfinite(in.uv[2 * i]) || !std::isfinite(in.uv[2 * i + 1]) ||
                !std::isfinite(in.weight[i]) || !(in.weight[i] > 0.0))
                continue;
            if (first < 0)
                first = in.obsCam[i];
            else if (in.obsCam[i] != first)
                return true;
        }
        return false;
    }

    // Control-point targets that count for the gauge (on tracks usable for it).
    std::vector<const Input::Constraint*> gaugeControls(const Input& in)
    {
        std::vector<const Input::Constraint*> out;
        for (const Input::Constraint& k : in.cons)
            if (!k.plane && trackUsableForGauge(in, k.track)) out.push_back(&k);
        return out;
    }

    // Three targets spanning a triangle fix pose and scale (hasCompleteControlPointGauge, :181-240).
    bool controlGaugeComplete(const std::vector<const Input::Constraint*>& ks)
    {
        const size_t n = ks.size();
        for (size_t a = 0; a < n; ++a)
            for (size_t b = a + 1; b < n; ++b)
            {
                double u[3], u2 = 0.0;
                for (int i = 0; i < 3; ++i)
                {
                    u[i] = ks[b]->target[i] - ks[a]->target[i];
                    u2 += u[i] * u[i];
                }
                if (u2 <= 1e-16) continue;
                for (size_t c = b + 1; c < n; ++c)
                {
                    double v[3], v2 = 0.0;
                    for (int i = 0; i < 3; ++i)
                    {
                        v[i] = ks[c]->target[i] - ks[a]->target[i];
                        v2 += v[i] * v[i];
                    }
                    const double x0 = u[1] * v[2] - u[2] * v[1], x1 = u[2] * v[0] - u[0] * v[2],
                                 x2 = u[0] * v[1] - u[1] * v[0];
                    if (v2 > 1e-16 && x0 * x0 + x1 * x1 + x2 * x2 > 1e-12 * u2 * v2) return true;
                }
            }
        return false;
    }

    // A target away from a fixed camera's centre gives the scale (hasControlPointBaselineToFixedCamera, :147-179).
    bool controlScaleFromFixed(const Input& in, const std::vector<const Input::Constraint*>& ks, const std::set<int>& fixed)
    {
        for (const Input::Constraint* k : ks)
            for (int c : fixed)
            {
                const double* C = &in.C[3 * c];
                if (!std::isfinite(C[0]) || !std::isfinite(C[1]) || !std::isfinite(C[2])) continue;
                double d = 0.0;
                for (int i = 0; i < 3; ++i) d += (k->target[i] - C[i]) * (k->target[i] - C[i]);
                if (d > 1e-16) return true;
            }
        return false;
    }

    std::string normalizeGauge(const Input& in, const pb_solve_options& o, Prepared* p)
    {
        const int n = in.nCam;
        std::set<int> fixed;
        for (int c : in.fixedCams)
        {
            if (c < 0 || c >= n) return "invalid: fixed camera index out of range";
            if (!fixed.insert(c).second) return "invalid: duplicate fixed camera index";
        }
        const bool requested = in.gaugeAnchor >= 0 || in.gaugeScale >= 0 || in.gaugeBaseline > 0.0;
        if (requested && (in.gaugeAnchor < 0 || in.gaugeScale < 0 || in.gaugeAnchor >= n || in.gaugeScale >= n ||
                          in.gaugeAnchor == in.gaugeScale || !std::isfinite(in.gaugeBaseline) ||
                          in.gaugeBaseline <= 0.0))
            return "invalid: reference gauge";
        std::set<int> ft;
        for (int t : in.fixedTracks)
        {
            if (t < 0 || t >= in.nTracks) return "invalid: fixed track index out of range";
            if (!ft.insert(t).second) return "invalid: duplicate fixed track index";
        }
        p->fixedCams = in.fixedCams;
        p->fixedTracks.assign(ft.begin(), ft.end());
        p->anchor = in.gaugeAnchor;
        p->scaleCam = in.gaugeScale;
        p->baseline = in.gaugeBaseline;
        if (!o.refine_camera_pose || in.gaugePolicy == PB_GAUGE_CALLER_MANAGED || n == 0) return {};
        // control points: three non-collinear targets give absolute pose and scale; a target away from a fixed
        // camera's centre gives absolute scale (BundleAdjustValidation.cpp:826-850)
        const std::vector<const Input::Constraint*> ks = gaugeControls(in);
        const bool absolutePose = controlGaugeComplete(ks);
        bool absoluteScale = absolutePose || controlScaleFromFixed(in, ks, fixed);
        if (in.gaugePolicy == PB_GAUGE_REQUIRE_EXPLICIT)
            return (absolutePose && absoluteScale) || fixed.size() >= 2 || (fixed.size() == 1 && absoluteScale)
                       ? std::string()
                       : std::string("unsupported: incomplete gauge (fix two cameras, or one with an absolute scale)");
        if (absolutePose && absoluteScale) return {};
        if (fixed.empty())
        {
            p->fixedCams.push_back(0);
            fixed.insert(0);
            absoluteScale = absoluteScale || controlScaleFromFixed(in, ks, fixed);
        }
        if (!absoluteScale && fixed.size() < 2)
        {
            const int a = *fixed.begin();
            int best = -1;
            double bestD = 1e-20;
            for (int c = 0; c < n; ++c)
            {
                if (fixed.count(c)) continue;
                double d = 0.0;
                for (int k = 0; k < 3; ++k) d += (in.C[3 * c + k] - in.C[3 * a + k]) * (in.C[3 * c + k] - in.C[3 * a + k]);
                if (std::isfinite(d) && d > bestD)
                {
                    bestD = d;
                    best = c;
                }
            }
            if (best < 0) return "unsupported: no second camera with a non-zero baseline";
            p->anchor = a;
            p->scaleCam = best;
            p->baseline = std::sqrt(bestD);
        }
        return {};
    }

    void prepareActive(const Input& in, const pb_solve_options& o, Prepared* p)
    {
        const int n = in.nCam;
        p->activeTrack.assign(in.nTracks, 0);
        p->cameraBlock.assign(n, -1);
        p->intrinsicBlock.assign(n, -1);
        p->groupOfCam.assign(n, 0);
        p->trackBlock.assign(in.nTracks, -1);
        std::vector<char> hasResidual(n, 0);
        std::vector<pbdev::Camera> cams(n);
        for (int c = 0; c < n; ++c) cams[c] = inputCamera(in, c);
        for (int t = 0; t < in.nTracks; ++t)
        {
            const double* X = &in.X[3 * t];
            if (!std::isfinite(X[0]) || !std::isfinite(X[1]) || !std::isfinite(X[2])) continue;
            int first = -1, count = 0;
            bool second = false, front = true;
            double sq = 0.0;
            for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
            {
                if (!usable(in, i)) continue;
                const int c = in.obsCam[i];
                first = first < 0 ? c : first;
                second = second || c != first;
                double u, v;
                if (!pbdev::project(cams[c], X[0], X[1], X[2], u, v))
                {
                    front = false;
                    break;
                }
                const double du = u - in.uv[2 * i], dv = v - in.uv[2 * i + 1];
                sq += sanitizedWeight(in.weight[i]) * (du * du + dv * dv);
                ++count;
            }
            if (count < 2 || !second || !front) continue;
            const double rms = std::sqrt(sq / (2.0 * count));
            if (o.max_initial_track_rms > 0.0 && rms > o.max_initial_track_rms)
            {
                ++p->rejectedInitial;
                continue;
            }
            p->activeTrack[t] = 1;
            ++p->activeTracks;
            p->observations += count;
            for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
                if (usable(in, i)) hasResidual[in.obsCam[i]] = 1;
        }
        p->rig = in.rig();
        if (o.refine_camera_pose && !p->rig)
        {
            std::set<int> fixed(p->fixedCams.begin(), p->fixedCams.end());
            for (int c = 0; c < n; ++c)
                if (hasResidual[c] && !fixed.count(c)) p->cameraBlock[c] = p->cameraBlocks++;
        }
        else if (p->rig)
        {
            // a capture is refined when one of its cameras has a residual and neither the capture, nor one of its
            // cameras is fixed, nor one of them is the gauge scale camera; a sensor when one of its cameras has a
            // residual and its extrinsic is not fixed; sensor blocks follow the capture blocks
            std::set<int> fixed(p->fixedCams.begin(), p->fixedCams.end());
            p->rigCapBlock.assign(in.nCap, -1);
            p->rigSenBlock.assign(in.nSen, -1);
            p->rigCapBlockByCam.assign(n, -1);
            p->rigSenBlockByCam.assign(n, -1);
            if (o.refine_camera_pose)
            {
                std::vector<char> capResidual(in.nCap, 0), capFixed(in.nCap, 0), senResidual(in.nSen, 0);
                for (int k = 0; k < in.nCap; ++k) capFixed[k] = in.capFixed[k] != 0;
                for (int c = 0; c < n; ++c)
                {
                    const int k = in.camCapture[c], q = in.camSensor[c];
                    capResidual[k] = capResidual[k] || hasResidual[c];
                    capFixed[k] = capFixed[k] || fixed.count(c) || p->scaleCam == c;
                    senResidual[q] = senResidual[q] || hasResidual[c];
                }
                for (int k = 0; k < in.nCap; ++k)
                    if (capResidual[k] && !capFixed[k]) p->rigCapBlock[k] = p->rigCapBlocks++;
                for (int q = 0; q < in.nSen; ++q)
                    if (senResidual[q] && !in.senFixed[q]) p->rigSenBlock[q] = p->rigCapBlocks + p->rigSenBlocks++;
                for (int c = 0; c < n; ++c)
                {
                    p->rigCapBlockByCam[c] = p->rigCapBlock[in.camCapture[c]];
                    p->rigSenBlockByCam[c] = p->rigSenBlock[in.camSensor[c]];
                    if (hasResidual[c] && (p->rigCapBlockByCam[c] >= 0 || p->rigSenBlockByCam[c] >= 0))
                        p->cameraBlock[c] = p->cameraBlocks++;
                }
            }
        }
        p->primaryBlocks = p->refCameraBlocks();
        if (o.shared_intrinsic_mask & 0x1FFFu)
        {
            std::vector<int> ids;
            for (int c = 0; c < n; ++c)
            {
                const int id = in.hasGroup ? in.group[c] : 0;
                auto it = std::find(ids.begin(), ids.end(), id);
                if (it == ids.end())
                {
                    p->groupOfCam[c] = static_cast<int>(ids.size());
                    ids.push_back(id);
                }
                else
                {
                    p->groupOfCam[c] = static_cast<int>(it - ids.begin());
                }
            }
            p->intrinsicBlocks = static_cast<int>(ids.size());
            for (int c = 0; c < n; ++c) p->intrinsicBlock[c] = p->primaryBlocks + p->groupOfCam[c];
            p->primaryBlocks += p->intrinsicBlocks;
        }
        for (int t = 0; t < in.nTracks; ++t)
        {
            if (!p->activeTrack[t]) continue;
            if (std::binary_search(p->fixedTracks.begin(), p->fixedTracks.end(), t)) continue;
            p->trackBlock[t] = p->trackBlocks++;
        }
    }

    double median(std::vector<double> v, double fallback)
    {
        v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return !std::isfinite(x); }), v.end());
        if (v.empty()) return fallback;
        std::sort(v.begin(), v.end());
        const size_t k = v.size();
        return k % 2 ? v[k / 2] : 0.5 * (v[k / 2 - 1] + v[k / 2]);
    }

    void prepareGroups(const Input& in, const pb_solve_options& o, Prepared* p)
    {
        const int G = p->intrinsicBlocks;
        p->G = G;
        p->gParams.assign(13 * G, 0.0);
        p->gPrior.assign(13 * G, 0.0);
        p->gInvSigma.assign(13 * G, 0.0);
        p->gLower.assign(13 * G, -kInf);
        p->gUpper.assign(13 * G, kInf);
        p->gEnabled.assign(G, 0u);
        p->gActive.assign(pbdev::kMaxActive * G, 0);
        p->gK.assign(G, 0);
        p->gUnit.assign(G, -1);
        p->slotToCompact.assign(13 * G, -1);
        const double low = std::max(1.0, o.low_order_distortion_scale);
        for (int g = 0; g < G; ++g)
        {
            std::array<std::vector<double>, 13> init, ref;
            std::vector<double> widths, heights;
            uint32_t sup = 0x1FFFu;
            for (int c = 0; c < in.nCam; ++c)
            {
                if (p->groupOfCam[c] != g) continue;
                const double* s = &in.slots[13 * c];
                const double* r = &in.ref[13 * c];
                init[0].push_back(s[0]);
                ref[0].push_back(r[0]);
                init[1].push_back(std::exp(s[1]));
                ref[1].push_back(std::exp(r[1]));
                init[2].push_back(s[2] - r[2]);
                init[3].push_back(s[3] - r[3]);
                ref[2].push_back(0.0);
                ref[3].push_back(0.0);
                for (int k = 4; k < 13; ++k)
                {
                    init[k].push_back(s[k]);
                    ref[k].push_back(r[k]);
                }
                sup &= supportedMask(in.conv[c]);
                // image size of the camera, else twice the reference principal point (Model.cpp:61-66)
                const bool sized = !in.imageSize.empty();
                widths.push_back(sized ? double(in.imageSize[2 * c]) : 2.0 * r[2]);
                heights.push_back(sized ? double(in.imageSize[2 * c + 1]) : 2.0 * r[3]);
            }
            double* gp = &p->gParams[13 * g];
            double* pr = &p->gPrior[13 * g];
            double* lo = &p->gLower[13 * g];
            double* hi = &p->gUpper[13 * g];
            const uint32_t en = o.shared_intrinsic_mask & sup & 0x1FFFu;
            p->gEnabled[g] = en;
            const double fref = median(ref[0], 1.0);
            const double aref = median(ref[1], 1.0);
            gp[0] = std::max(1e-12, median(init[0], fref));
            gp[1] = std::log(std::max(1e-12, median(init[1], aref)));
            pr[0] = fref;
            pr[1] = std::log(aref);
            for (int k = 2; k < 13; ++k)
            {
                pr[k] = median(ref[k], 0.0);
                gp[k] = median(init[k], pr[k]);
            }
            lo[0] = fref * o.min_focal_scale;
            hi[0] = fref * o.max_focal_scale;
            lo[1] = std::log(aref * o.min_aspect_scale);
            hi[1] = std::log(aref * o.max_aspect_scale);
            const double pl = fref * o.max_principal_offset_fraction;
            lo[2] = lo[3] = -pl;
            hi[2] = hi[3] = pl;
            const double lim[9] = {o.max_abs[0] * low, o.max_abs[1], o.max_abs[2], o.max_abs[3] * low,
                                   o.max_abs[4] * low, fref * o.max_abs[5], o.max_abs[6], o.max_abs[7], o.max_abs[8]};
            for (int k = 0; k < 9; ++k)
            {
                lo[4 + k] = -lim[k];
                hi[4 + k] = lim[k];
            }
            const double ps = std::max(1e-6, fref * o.principal_prior_sigma_fraction);
            const double sig[13] = {o.focal_prior_sigma * fref, o.aspect_prior_sigma, ps, ps,
                                    o.prior_sigma[0] * low, o.prior_sigma[1], o.prior_sigma[2],
                                    o.prior_sigma[3] * low, o.prior_sigma[4] * low, fref * o.prior_sigma[5],
                                    o.prior_sigma[6], o.prior_sigma[7], o.prior_sigma[8]};
            int K = 0;
            for (int k = 0; k < 13; ++k)
            {
                gp[k] = std::clamp(gp[k], lo[k], hi[k]);
                const bool on = (en >> k) & 1u;
                p->gInvSigma[13 * g + k] = on ? 1.0 / std::max(1e-9, sig[k]) : 0.0;
                if (on)
                {
                    if (K >= pbdev::kMaxActive) throw std::invalid_argument("unsupported: more refined slots than the 13-slot block");
                    p->gActive[pbdev::kMaxActive * g + K] = static_cast<int8_t>(k);
                    p->slotToCompact[13 * g + k] = K;
                    ++K;
                }
            }
            p->gK[g] = K;
            if (o.use_transition_prior)
            {
                // reference calibration transition prior (Model.cpp:173-225): every slot but the aspect trades
                // its prior sigma for a transition weight, which is non-zero only when the refined set differs
                // from the previous stage's; its scale follows the image half-diagonal r in normalised
                // coordinates (slot effect at the image corner), times 1e6 for a slot refined before, 0.01 else
                bool changed = false;
                for (int k = 0; k < 13; ++k)
                    if (k != 1) changed = changed || (((en >> k) & 1u) != ((o.previous_intrinsic_mask >> k) & 1u));
                const double nx = 0.5 * median(widths, 0.0) / fref;
                const double ny = 0.5 * median(heights, 0.0) / (fref * aref);
                const double r = std::sqrt(nx * nx + ny * ny);
                double tw[13] = {};
                if (changed && std::isfinite(r) && r > 0.0)
                {
                    const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;
                    const double base[13] = {1.0 / r, 0.0, 1.0, 1.0, 1.0 / (r3 * fref), 1.0 / (r5 * fref),
                                             1.0 / (r7 * fref), 1.0 / (r2 * fref), 1.0 / (r2 * fref), 1.0 / r,
                                             1.0 / (r9 * fref), 1.0, 1.0};
                    for (int k = 0; k < 13; ++k)
                        if (k != 1 && ((en >> k) & 1u))
                            tw[k] = 1.0 / (base[k] * (((o.previous_intrinsic_mask >> k) & 1u) ? 1e6 : 0.01));
                }
                for (int k = 0; k < 13; ++k)
                    if (k != 1) p->gInvSigma[13 * g + k] = tw[k];
            }
        }
        // units: pose units first (ids = camera block ids), then ceil(K/6) per group
        int u = p->cameraBlocks;
        p->unitWidth.assign(p->cameraBlocks, 6);
        p->globalDim = 6 * p->cameraBlocks;
        for (int g = 0; g < G; ++g)
        {
            if (p->gK[g] == 0) continue;
            p->gUnit[g] = u;
            const int m = (p->gK[g] + 5) / 6;
            for (int j = 0; j < m; ++j) p->unitWidth.push_back(std::min(6, p->gK[g] - 6 * j));
            u += m;
            p->globalDim += p->gK[g];
        }
        p->nUnits = u;
    }

    // ------------------------------------------------------------------------------------------
    // Device layout of one prepared problem.
    // Block-sparse rows of a reduced system (both triangles) and its fused-PCG wave slots.
    struct Rows
    {
        std::vector<int32_t> rowPtr, colIdx, diagPos, blockPosUp, blockPosLo;
        std::vector<int32_t> cgWaveRow, cgWaveBegin, cgWaveEnd, cgWaveLead;
        int nNnz = 0;
    };

    struct Layout
    {
        int nActive = 0, nObs = 0, nElim = 0, nEntries = 0, nBlocks = 0, nNnz = 0, nItems = 0;
        long nPairs = 0;
        int maxGroupK = 0;
        std::vector<int32_t> activeTrackOf;  // active index -> input track
        std::vector<int32_t> activeIndexOf;  // input track -> active index or -1
        std::vector<int64_t> obsInput;       // active observation -> input observation index
        std::vector<int32_t> obsOffset, obsCam, obsPoint, obsPoseEntry, obsGroupEntry;
        std::vector<double> obsUV, obsW, pointX;
        std::vector<int32_t> pointElim, elimTrack;
        std::vector<int32_t> camObsOffset, camObs, camUnit, camGroup, camDiagBlock, camGroupBlock;
        std::vector<int32_t> groupCamOffset, groupCams, groupPairBlock;
        std::vector<int32_t> entryOffset, entryUnit, entryPoint;
        std::vector<int32_t> blockRow, blockCol, blockItemOffset;
        std::vector<int32_t> itemBlock, itemBegin, pairE1, pairE2;
        Rows rows;  // block-sparse rows of S (both triangles) and the fused-PCG wave slots
        // camera rig: the rig system (units: refined captures, refined sensors, the intrinsic-group units), its
        // upper blocks with the camera-level contributions of each, the right-hand-side and expansion maps
        int nSysUnits = 0, sysGlobalDim = 0;
        std::vector<int32_t> sysUnitWidth, sysBlockRow, sysBlockCol;
        Rows sysRows;
        std::vector<int32_t> projOffset, projPos, projMapL, projMapR;
        std::vector<int32_t> sysRhsOffset, sysRhsUnit, sysRhsMap, expUnit, expMap;
        // lane-per-observation point linearisation: observation and eliminated point of each lane of each wave
        // (-1 past a wave's last point); the points it cannot take go to the thread-per-point kernel
        std::vector<int32_t> linLaneObs, linLanePoint, linFallback;
    };

    // Block-sparse rows of a reduced system (both triangles) with the fused-PCG wave slots, from its upper
    // blocks sorted by (row, column); blockStart[u] is the first upper block of row u.
    void blockRows(int U, const std::vector<int32_t>& blockRow, const std::vector<int32_t>& blockCol,
                   const std::vector<int32_t>& blockStart, Rows* R)
    {
        const int nBlocks = static_cast<int>(blockRow.size());
        // full block-sparse rows: lower part (from upper blocks (w, u), w < u) then the upper part
        std::vector<std::vector<int32_t>> lower(U);
        for (int blk = 0; blk < nBlocks; ++blk)
            if (blockRow[blk] != blockCol[blk]) lower[blockCol[blk]].push_back(blk);
        R->rowPtr.assign(U + 1, 0);
        R->blockPosUp.assign(nBlocks, -1);
        R->blockPosLo.assign(nBlocks, -1);
        R->diagPos.assign(U, -1);
        R->colIdx.clear();
        for (int u = 0; u < U; ++u)
        {
            for (int blk : lower[u])
            {
                R->blockPosLo[blk] = static_cast<int32_t>(R->colIdx.size());
                R->colIdx.push_back(blockRow[blk]);
            }
            for (int blk = blockStart[u]; blk < blockStart[u + 1]; ++blk)
            {
                R->blockPosUp[blk] = static_cast<int32_t>(R->colIdx.size());
                if (blockCol[blk] == u)
                {
                    R->diagPos[u] = R->blockPosUp[blk];
                    R->blockPosLo[blk] = R->blockPosUp[blk];
                }
                R->colIdx.push_back(blockCol[blk]);
            }
            R->rowPtr[u + 1] = static_cast<int32_t>(R->colIdx.size());
        }
        R->nNnz = static_cast<int>(R->colIdx.size());
        // fused-PCG wave slots: a row gets one wave per kCgSegment stored blocks (at most kCgWaves, all in
        // one workgroup); rows are packed into workgroups of kCgWaves slots in row order
        R->cgWaveRow.clear();
        R->cgWaveBegin.clear();
        R->cgWaveEnd.clear();
        R->cgWaveLead.clear();
        auto padPack = [&]() {
            while (R->cgWaveRow.size() % pbk::kCgWaves)
            {
                R->cgWaveRow.push_back(-1);
                R->cgWaveBegin.push_back(0);
                R->cgWaveEnd.push_back(0);
                R->cgWaveLead.push_back(0);
            }
        };
        for (int u = 0; u < U; ++u)
        {
            const int b0 = R->rowPtr[u], n = R->rowPtr[u + 1] - b0;
            const int w = std::max(1, std::min(pbk::kCgWaves, (n + pbk::kCgSegment - 1) / pbk::kCgSegment));
            const int len = (n + w - 1) / w;
            const int used = static_cast<int>(R->cgWaveRow.size() % pbk::kCgWaves);
            if (used + w > pbk::kCgWaves) padPack();
            for (int j = 0; j < w; ++j)
            {
                R->cgWaveRow.push_back(u);
                R->cgWaveBegin.push_back(b0 + std::min(n, j * len));
                R->cgWaveEnd.push_back(b0 + std::min(n, (j + 1) * len));
                R->cgWaveLead.push_back(j == 0 ? w : 0);
            }
        }
        padPack();
        // the kernel relies on these: every row owned by exactly one lead slot whose row slots follow it in the
        // same workgroup, and slot ranges inside the row
        std::vector<int> owners(U, 0);
        const int nSlots = static_cast<int>(R->cgWaveRow.size());
        for (int sl = 0; sl < nSlots; ++sl)
        {
            const int u = R->cgWaveRow[sl], lead = R->cgWaveLead[sl];
            if (u < 0) continue;
            if (u >= U || R->cgWaveBegin[sl] < R->rowPtr[u] || R->cgWaveEnd[sl] > R->rowPtr[u + 1] ||
                R->cgWaveBegin[sl] > R->cgWaveEnd[sl])
                throw std::logic_error("PCG wave slot range");
            if (lead > 0)
            {
                ++owners[u];
                if (sl % pbk::kCgWaves + lead > pbk::kCgWaves) throw std::logic_error("PCG row crosses workgroup");
                for (int j = 1; j < lead; ++j)
                    if (R->cgWaveRow[sl + j] != u || R->cgWaveLead[sl + j] != 0)
                        throw std::logic_error("PCG row slots not contiguous");
            }
        }
        for (int u = 0; u < U; ++u)
            if (owners[u] != 1) throw std::logic_error("PCG row without a unique owner");
    }

    // The rig system of a camera rig problem. A camera-level pose unit a (camera c) maps to its capture's rig
    // unit through chain map 2c and to its sensor's through 2c + 1 (k_rig_maps); a group unit maps to its rig
    // unit unchanged (-1). Every stored camera-level block (a, b) adds M_l^T S_ab M_r to the rig block (u, v),
    // u <= v, of each pair of its images; the contributions of a rig block are kept in camera-level storage
    // order, so every build sums them in the same order.
    void buildSysLayout(const Input& in, const Prepared& p, Layout* L)
    {
        const int rigPose = p.rigCapBlocks + p.rigSenBlocks;
        const int U = p.nUnits, V = rigPose + (p.nUnits - p.cameraBlocks);
        L->nSysUnits = V;
        L->sysUnitWidth.assign(rigPose, 6);
        for (int a = p.cameraBlocks; a < U; ++a) L->sysUnitWidth.push_back(p.unitWidth[a]);
        L->sysGlobalDim = 6 * rigPose;
        for (int g = 0; g < p.G; ++g) L->sysGlobalDim += p.gK[g];
        std::vector<int> camOfUnit(p.cameraBlocks, -1);
        for (int c = 0; c < in.nCam; ++c)
            if (p.cameraBlock[c] >= 0) camOfUnit[p.cameraBlock[c]] = c;
        L->expUnit.assign(2 * size_t(U), -1);
        L->expMap.assign(2 * size_t(U), -1);
        for (int a = 0; a < U; ++a)
        {
            if (a < p.cameraBlocks)
            {
                const int c = camOfUnit[a];
                int e = 0;
                if (p.rigCapBlockByCam[c] >= 0)
                {
                    L->expUnit[2 * a + e] = p.rigCapBlockByCam[c];
                    L->expMap[2 * a + e++] = 2 * c;
                }
                if (p.rigSenBlockByCam[c] >= 0)
                {
                    L->expUnit[2 * a + e] = p.rigSenBlockByCam[c];
                    L->expMap[2 * a + e++] = 2 * c + 1;
                }
            }
            else
            {
                L->expUnit[2 * a] = a - p.cameraBlocks + rigPose;
            }
        }
        // right-hand side: the camera-level units of each rig unit, in unit order
        std::vector<std::vector<std::pair<int32_t, int32_t>>> src(V);
        for (int a = 0; a < U; ++a)
            for (int e = 0; e < 2; ++e)
                if (L->expUnit[2 * a + e] >= 0) src[L->expUnit[2 * a + e]].push_back({a, L->expMap[2 * a + e]});
        L->sysRhsOffset.assign(V + 1, 0);
        for (int u = 0; u < V; ++u)
        {
            for (const auto& [a, m] : src[u])
            {
                L->sysRhsUnit.push_back(a);
                L->sysRhsMap.push_back(m);
            }
            L->sysRhsOffset[u + 1] = static_cast<int32_t>(L->sysRhsUnit.size());
        }
        // projection: contributions keyed by rig upper block
        struct Contribution
        {
            int32_t u, v, pos, ml, mr;
        };
        std::vector<Contribution> cs;
        for (int a = 0; a < U; ++a)
            for (int k = L->rows.rowPtr[a]; k < L->rows.rowPtr[a + 1]; ++k)
            {
                const int b = L->rows.colIdx[k];
                for (int ea = 0; ea < 2; ++ea)
                {
                    const int u = L->expUnit[2 * a + ea];
                    if (u < 0) continue;
                    for (int eb = 0; eb < 2; ++eb)
                    {
                        const int v = L->expUnit[2 * b + eb];
                        if (v < 0 || u > v) continue;
                        cs.push_back({u, v, k, L->expMap[2 * a + ea], L->expMap[2 * b + eb]});
                    }
                }
            }
        std::stable_sort(cs.begin(), cs.end(),
                         [](const Contribution& x, const Contribution& y) { return x.u != y.u ? x.u < y.u : x.v < y.v; });
        std::vector<int32_t> blockStart(V + 1, 0);
        L->projOffset.assign(1, 0);
        for (size_t i = 0; i < cs.size(); ++i)
        {
            if (i == 0 || cs[i].u != cs[i - 1].u || cs[i].v != cs[i - 1].v)
            {
                if (i > 0) L->projOffset.push_back(static_cast<int32_t>(i));
                L->sysBlockRow.push_back(cs[i].u);
                L->sysBlockCol.push_back(cs[i].v);
            }
            L->projPos.push_back(cs[i].pos);
            L->projMapL.push_back(cs[i].ml);
            L->projMapR.push_back(cs[i].mr);
        }
        if (!cs.empty()) L->projOffset.push_back(static_cast<int32_t>(cs.size()));
        // every rig unit keeps its diagonal block (an empty one finalizes to the identity)
        std::vector<int32_t> row = L->sysBlockRow, col = L->sysBlockCol, off = L->projOffset;
        L->sysBlockRow.clear();
        L->sysBlockCol.clear();
        L->projOffset.assign(1, 0);
        std::vector<int32_t> pos, ml, mr;
        size_t b = 0;
        for (int u = 0; u < V; ++u)
        {
            const bool hasDiag = b < row.size() && row[b] == u && col[b] == u;
            if (!hasDiag)
            {
                L->sysBlockRow.push_back(u);
                L->sysBlockCol.push_back(u);
                L->projOffset.push_back(static_cast<int32_t>(pos.size()));
            }
            for (; b < row.size() && row[b] == u; ++b)
            {
                L->sysBlockRow.push_back(row[b]);
                L->sysBlockCol.push_back(col[b]);
                for (int j = off[b]; j < off[b + 1]; ++j)
                {
                    pos.push_back(L->projPos[j]);
                    ml.push_back(L->projMapL[j]);
                    mr.push_back(L->projMapR[j]);
                }
                L->projOffset.push_back(static_cast<int32_t>(pos.size()));
            }
            blockStart[u + 1] = static_cast<int32_t>(L->sysBlockRow.size());
        }
        L->projPos.swap(pos);
        L->projMapL.swap(ml);
        L->projMapR.swap(mr);
        blockRows(V, L->sysBlockRow, L->sysBlockCol, blockStart, &L->sysRows);
    }

    void buildLayout(const Input& in, const Prepared& p, Layout* L)
    {
        // Active tracks and their usable observations. The device order of the active tracks (and with it of
        // the eliminated points, their cross-block entries and observations) is sorted by the lowest camera
        // that observes the track: cameras are numbered along the block, so the points shared by a pair of
        // cameras sit close together and the Schur pairs of one block read neighbouring entries. The order
        // is internal; block numbering and masks keep the reference's track order.
        std::vector<int32_t> minCam;
        for (int t = 0; t < in.nTracks; ++t)
            if (p.activeTrack[t])
            {
                int m = std::numeric_limits<int>::max();
                for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
                    if (usable(in, i)) m = std::min(m, in.obsCam[i]);
                L->activeTrackOf.push_back(t);
                minCam.push_back(m);
            }
        {
            std::vector<int32_t> order(L->activeTrackOf.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return minCam[x] < minCam[y]; });
            std::vector<int32_t> sorted(order.size());
            for (size_t i = 0; i < order.size(); ++i) sorted[i] = L->activeTrackOf[order[i]];
            L->activeTrackOf.swap(sorted);
        }
        L->nActive = static_cast<int>(L->activeTrackOf.size());
        L->activeIndexOf.assign(in.nTracks, -1);
        for (int a = 0; a < L->nActive; ++a) L->activeIndexOf[L->activeTrackOf[a]] = a;
        L->obsOffset.assign(L->nActive + 1, 0);
        for (int a = 0; a < L->nActive; ++a)
        {
            const int t = L->activeTrackOf[a];
            for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
            {
                if (!usable(in, i)) continue;
                L->obsInput.push_back(i);
                L->obsCam.push_back(in.obsCam[i]);
                L->obsPoint.push_back(a);
                L->obsUV.push_back(in.uv[2 * i]);
                L->obsUV.push_back(in.uv[2 * i + 1]);
                const double ms = sanitizedScale(in.scale[i]);
                L->obsW.push_back(sanitizedWeight(in.weight[i]) / (ms * ms));
            }
            L->obsOffset[a + 1] = static_cast<int32_t>(L->obsCam.size());
            L->pointX.insert(L->pointX.end(), &in.X[3 * t], &in.X[3 * t] + 3);
        }
        L->nObs = static_cast<int>(L->obsCam.size());
        L->pointElim.assign(L->nActive, -1);
        for (int a = 0; a < L->nActive; ++a)
            if (p.trackBlock[L->activeTrackOf[a]] >= 0)
            {
                L->pointElim[a] = static_cast<int32_t>(L->elimTrack.size());
                L->elimTrack.push_back(a);
            }
        L->nElim = static_cast<int>(L->elimTrack.size());
        // camera units and groups
        const int n = in.nCam;
        L->camUnit.assign(n, -1);
        L->camGroup.assign(n, -1);
        for (int c = 0; c < n; ++c)
        {
            L->camUnit[c] = p.cameraBlock[c];
            if (p.intrinsicBlock[c] >= 0) L->camGroup[c] = p.groupOfCam[c];
        }
        for (int g = 0; g < p.G; ++g) L->maxGroupK = std::max(L->maxGroupK, p.gK[g]);
        auto groupUnits = [&](int c, int* first) -> int
        {
            const int g = L->camGroup[c];
            if (g < 0 || p.gK[g] == 0) return 0;
            *first = p.gUnit[g];
            return (p.gK[g] + 5) / 6;
        };
        // camera-sorted observation permutation
        L->camObsOffset.assign(n + 1, 0);
        for (int o = 0; o < L->nObs; ++o) ++L->camObsOffset[L->obsCam[o] + 1];
        for (int c = 0; c < n; ++c) L->camObsOffset[c + 1] += L->camObsOffset[c];
        L->camObs.assign(L->nObs, 0);
        {
            std::vector<int32_t> fill(L->camObsOffset.begin(), L->camObsOffset.end() - 1);
            for (int o = 0; o < L->nObs; ++o) L->camObs[fill[L->obsCam[o]]++] = o;
        }
        // cross-block entries per eliminated point
        L->obsPoseEntry.assign(L->nObs, -1);
        L->obsGroupEntry.assign(L->nObs, -1);
        L->entryOffset.assign(L->nElim + 1, 0);
        std::vector<int32_t> units;
        for (int b = 0; b < L->nElim; ++b)
        {
            const int a = L->elimTrack[b];
            units.clear();
            for (int o = L->obsOffset[a]; o < L->obsOffset[a + 1]; ++o)
            {
                const int c = L->obsCam[o];
                if (L->camUnit[c] >= 0) units.push_back(L->camUnit[c]);
                int g0 = 0;
                const int m = groupUnits(c, &g0);
                for (int j = 0; j < m; ++j) units.push_back(g0 + j);
            }
            std::sort(units.begin(), units.end());
            units.erase(std::unique(units.begin(), units.end()), units.end());
            const int base = static_cast<int>(L->entryUnit.size());
            for (int u : units)
            {
                L->entryUnit.push_back(u);
                L->entryPoint.push_back(b);
            }
            L->entryOffset[b + 1] = static_cast<int32_t>(L->entryUnit.size());
            for (int o = L->obsOffset[a]; o < L->obsOffset[a + 1]; ++o)
            {
                const int c = L->obsCam[o];
                if (L->camUnit[c] >= 0)
                    L->obsPoseEntry[o] =
                        base + static_cast<int>(std::lower_bound(units.begin(), units.end(), L->camUnit[c]) - units.begin());
                int g0 = 0;
                if (groupUnits(c, &g0) > 0)
                    L->obsGroupEntry[o] =
                        base + static_cast<int>(std::lower_bound(units.begin(), units.end(), g0) - units.begin());
            }
        }
        L->nEntries = static_cast<int>(L->entryUnit.size());
        // Points packed whole, in order, into waves of 64 observation lanes. A point with no observation, with
        // more than 64, or with two observations in one camera (two lanes would write one cross block) goes to the
        // thread-per-point kernel.
        {
            int fill = pbk::kWaveLanes;
            std::vector<int32_t> poses;
            for (int b = 0; b < L->nElim; ++b)
            {
                const int a = L->elimTrack[b];
                const int o0 = L->obsOffset[a], o1 = L->obsOffset[a + 1], n = o1 - o0;
                poses.clear();
                for (int o = o0; o < o1; ++o)
                    if (L->obsPoseEntry[o] >= 0) poses.push_back(L->obsPoseEntry[o]);
                std::sort(poses.begin(), poses.end());
                if (n == 0 || n > pbk::kWaveLanes || std::adjacent_find(poses.begin(), poses.end()) != poses.end())
                {
                    L->linFallback.push_back(b);
                    continue;
                }
                if (fill + n > pbk::kWaveLanes)
                {
                    L->linLaneObs.insert(L->linLaneObs.end(), pbk::kWaveLanes, -1);
                    L->linLanePoint.insert(L->linLanePoint.end(), pbk::kWaveLanes, -1);
                    fill = 0;
                }
                const size_t base = L->linLaneObs.size() - pbk::kWaveLanes;
                for (int o = o0; o < o1; ++o, ++fill)
                {
                    L->linLaneObs[base + fill] = o;
                    L->linLanePoint[base + fill] = b;
                }
            }
        }
        // Schur pairs bucketed by row unit, then sorted by (column unit, first entry)
        const int U = p.nUnits;
        std::vector<int64_t> rowCount(U + 1, 0);
        for (int b = 0; b < L->nElim; ++b)
        {
            const int e0 = L->entryOffset[b], e1 = L->entryOffset[b + 1];
            for (int i = e0; i < e1; ++i) rowCount[L->entryUnit[i] + 1] += e1 - i;
        }
        for (int u = 0; u < U; ++u) rowCount[u + 1] += rowCount[u];
        L->nPairs = rowCount[U];
        if (L->nPairs > std::numeric_limits<int32_t>::max()) throw std::invalid_argument("unsupported: too many Schur pairs");
        struct PairRec
        {
            int32_t col, e1, e2;
        };
        std::vector<PairRec> recs(L->nPairs);
        {
            std::vector<int64_t> fill(rowCount.begin(), rowCount.end() - 1);
            for (int b = 0; b < L->nElim; ++b)
            {
                const int e0 = L->entryOffset[b], e1 = L->entryOffset[b + 1];
                for (int i = e0; i < e1; ++i)
                    for (int j = i; j < e1; ++j)
                        recs[fill[L->entryUnit[i]]++] = {L->entryUnit[j], i, j};
            }
        }
        for (int u = 0; u < U; ++u)
            std::sort(recs.begin() + rowCount[u], recs.begin() + rowCount[u + 1],
                      [](const PairRec& x, const PairRec& y) { return x.col != y.col ? x.col < y.col : x.e1 < y.e1; });
        // direct-part blocks
        std::vector<std::vector<int32_t>> direct(U);
        for (int u = 0; u < U; ++u) direct[u].push_back(u);
        for (int c = 0; c < n; ++c)
        {
            const int pu = L->camUnit[c];
            int g0 = 0;
            const int m = groupUnits(c, &g0);
            if (pu >= 0)
                for (int j = 0; j < m; ++j) direct[pu].push_back(g0 + j);
        }
        for (int g = 0; g < p.G; ++g)
        {
            if (p.gK[g] == 0) continue;
            const int m = (p.gK[g] + 5) / 6;
            for (int i = 0; i < m; ++i)
                for (int j = i; j < m; ++j) direct[p.gUnit[g] + i].push_back(p.gUnit[g] + j);
        }
        // upper blocks in (row, col) order with their pair runs and work items
        std::vector<int32_t> blockStart(U + 1, 0);
        std::vector<int32_t> rowCols;
        L->itemBegin.clear();
        L->blockItemOffset.push_back(0);
        int64_t pairOut = 0;
        L->pairE1.resize(L->nPairs);
        L->pairE2.resize(L->nPairs);
        for (int u = 0; u < U; ++u)
        {
            rowCols.assign(direct[u].begin(), direct[u].end());
            for (int64_t k = rowCount[u]; k < rowCount[u + 1]; ++k)
                if (rowCols.empty() || rowCols.back() != recs[k].col) rowCols.push_back(recs[k].col);
            std::sort(rowCols.begin(), rowCols.end());
            rowCols.erase(std::unique(rowCols.begin(), rowCols.end()), rowCols.end());
            int64_t k = rowCount[u];
            for (int v : rowCols)
            {
                const int blk = static_cast<int>(L->blockRow.size());
                L->blockRow.push_back(u);
                L->blockCol.push_back(v);
                const int64_t start = pairOut;
                while (k < rowCount[u + 1] && recs[k].col == v)
                {
                    L->pairE1[pairOut] = recs[k].e1;
                    L->pairE2[pairOut] = recs[k].e2;
                    ++pairOut;
                    ++k;
                }
                for (int64_t s = start; s < pairOut; s += pbk::kPairChunk)
                {
                    L->itemBlock.push_back(blk);
                    L->itemBegin.push_back(static_cast<int32_t>(s));
                }
                L->blockItemOffset.push_back(static_cast<int32_t>(L->itemBlock.size()));
            }
            blockStart[u + 1] = static_cast<int32_t>(L->blockRow.size());
        }
        L->itemBegin.push_back(static_cast<int32_t>(pairOut));
        L->nItems = static_cast<int>(L->itemBlock.size());
        L->nBlocks = static_cast<int>(L->blockRow.size());
        recs.clear();
        recs.shrink_to_fit();
        auto blockId = [&](int u, int v) -> int
        {
            const auto b = L->blockCol.begin() + blockStart[u];
            const auto e = L->blockCol.begin() + blockStart[u + 1];
            const auto it = std::lower_bound(b, e, v);
            return (it != e && *it == v) ? static_cast<int>(it - L->blockCol.begin()) : -1;
        };
        blockRows(U, L->blockRow, L->blockCol, blockStart, &L->rows);
        L->nNnz = L->rows.nNnz;
        // direct-part block ids
        L->camDiagBlock.assign(n, -1);
        L->camGroupBlock.assign(size_t(pbk::kGroupUnits) * n, -1);
        for (int c = 0; c < n; ++c)
        {
            const int pu = L->camUnit[c];
            if (pu < 0) continue;
            L->camDiagBlock[c] = blockId(pu, pu);
            int g0 = 0;
            const int m = groupUnits(c, &g0);
            for (int j = 0; j < m; ++j) L->camGroupBlock[pbk::kGroupUnits * c + j] = blockId(pu, g0 + j);
        }
        L->groupPairBlock.assign(size_t(pbk::kGroupUnits) * pbk::kGroupUnits * std::max(1, p.G), -1);
        L->groupCamOffset.assign(p.G + 1, 0);
        for (int g = 0; g < p.G; ++g)
        {
            if (p.gK[g] > 0)
            {
                const int m = (p.gK[g] + 5) / 6;
                for (int i = 0; i < m; ++i)
                    for (int j = i; j < m; ++j)
                        L->groupPairBlock[(pbk::kGroupUnits * g + i) * pbk::kGroupUnits + j] =
                            blockId(p.gUnit[g] + i, p.gUnit[g] + j);
            }
            for (int c = 0; c < n; ++c)
                if (L->camGroup[c] == g) L->groupCams.push_back(c);
            L->groupCamOffset[g + 1] = static_cast<int32_t>(L->groupCams.size());
        }
    }

    // ------------------------------------------------------------------------------------------
    // Device buffers.
    struct DeviceAlloc
    {
        std::vector<void*> ptrs;
        size_t bytes = 0;
        template <typename V>
        V* alloc(size_t n)
        {
            void* p = nullptr;
            const size_t sz = std::max<size_t>(1, n) * sizeof(V);
            const hipError_t e = hipMalloc(&p, sz);
            if (e == hipErrorOutOfMemory) throw std::bad_alloc();
            check(e, "hipMalloc");
            ptrs.push_back(p);
            bytes += sz;
            return static_cast<V*>(p);
        }
        template <typename V>
        V* upload(const std::vector<V>& h, hipStream_t s)
        {
            V* d = alloc<V>(h.size());
            if (!h.empty()) check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, s), "upload");
            return d;
        }
        void release()
        {
            for (void* p : ptrs) (void)hipFree(p);
            ptrs.clear();
            bytes = 0;
        }
    };

    struct TimedSpan
    {
        hipEvent_t start, stop;
        int launches;
    };

    struct Family
    {
        int64_t launches = 0;
        double ms = 0.0;
        double bytes = 0.0;
        std::vector<TimedSpan> pending;
    };

    const char* const kFamilies[] = {"linearize_points", "camera_direct", "schur_blocks", "pcg_iteration", "cost",
                                     "back_substitution", "dense_factor", "dense_solve", "band_factor", "band_solve"};
    constexpr int kFamilyDenseFactor = 6, kFamilyDenseSolve = 7, kFamilyBandFactor = 8, kFamilyBandSolve = 9;
    // pinned readback area (in doubles): scalars of a build / step, flags, PCG scalars, PCG flags
    constexpr int kPinFlags = 32;
    constexpr int kPinPcg = 40;
    constexpr int kPinPcgFlags = 72;
    static_assert(pbk::kScalarCount <= kPinFlags && kPinPcg + pbk::kScalarCount <= kPinPcgFlags, "pinned layout");
    // PCG iterations are timed per chunk of back-to-back launches between two host polls (a pair of events
    // per chunk, so the events do not stretch the solve); chunks that contain the stop are not counted, so
    // every timed launch did a full iteration
    constexpr int kFamilyCount = sizeof(kFamilies) / sizeof(kFamilies[0]);
    constexpr int kPinDense = 80;       // dense solver words (8 int32) and two norms
    static_assert(kPinPcgFlags + pbk::kIflagCount / 2 <= kPinDense, "pinned layout");
    // a step enqueued with one host wait: the scalars after back-substitution and the first trial cost
    constexpr int kPinStep = 96, kPinTrial = kPinStep + pbk::kScalarCount;
    static_assert(kPinDense + 6 <= kPinStep && kPinTrial < 512, "pinned layout");
    // rows up to which the adjoint factors its reduced system densely (34 GB of FP64; FP64 PCG above that)
    constexpr long kDenseAdjointMax = 65536;
    // default largest reduced system solved densely per LM step (reduced_solver = auto): above it the
    // factorisation (n^3/3 flops per step) costs more than block-Jacobi PCG on the survey-sized blocks
    constexpr int kDenseSolveMaxRows = 16384;
    constexpr int kDenseSeedRefinements = 10;  // FP64 refinement steps of an FP32 factorisation before it is rejected
    // consecutive rejected FP32 seeds after which a solve factors in FP64 only (the system's conditioning is
    // beyond FP32 and changes slowly between LM steps)
    constexpr int kSeedRejectionsToStop = 2;
    // a mixed batch moves a window to FP64 assembly once an FP32 step lowers its objective by less than this
    // (relative): FP32 steps carry the early, large decreases; near the minimum the FP32 rounding of the
    // reduced system limits them, so the last iterations of every window assemble in FP64
    constexpr double kMixedSwitchDecrease = 1e-2;
    // a mixed single solve moves to FP64 assembly once an accepted FP32 step lowers the objective by less than this
    // (relative): about ten times the FP32 rounding of the assembled model, below which FP32 steps stop being
    // reliable descent steps along the weakly determined directions
    constexpr double kMixedSolveSwitchDecrease = 1e-6;

    // Host side of a batched solve: window ranges of the combined input and the per-window counts.
    struct BatchHost
    {
        int nWin = 0;
        std::vector<int32_t> camBase, trackBase;  // [nWin+1] input camera / track ranges of the windows
        std::vector<int32_t> gaugeAnchor, gaugeScale;  // per window after normalisation (window numbering), -1 = none
        std::vector<double> gaugeBaseline;
        std::vector<int32_t> observations, activeTracks, cameraBlocks, primaryBlocks;
        pbk::Batch B;
        double clockHz = 1e8;
    };

    // Launch shape of the dense window factorisation per assembly precision ([0] FP32, [1] FP64 assembly): every
    // usable shape is timed on the first round of a batch (each leaves bitwise the same result) and the fastest
    // is kept for the later rounds and for later batches of the same class: window count rounded up to a power
    // of two and largest window system rounded up to a multiple of 64 unknowns (the two figures that set the
    // occupancy and the longest factorisation).
    struct FactorTuning
    {
        std::vector<int32_t> key;
        int shape[2] = {-1, -1};   // -1: not timed yet
        double ms[2][pbk::kFactorShapes] = {};
        bool usable[2][pbk::kFactorShapes] = {};
        hipEvent_t ev[pbk::kFactorShapes][2] = {};
    };
}  // namespace

struct pb_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error = "ok";
    std::string deviceName;
    Input in;
    bool uploaded = false;
    // prepared structure (kept while the problem and the structural options stay the same)
    bool structureValid = false;
    pb_solve_options structureOptions{};  // options the prepared structure was built with
    Prepared prep;
    Layout lay;
    DeviceAlloc dstruct;  // structure + state
    DeviceAlloc dwork32, dwork64;
    DeviceAlloc dbatch;      // per-window controller state of a batched solve
    DeviceAlloc dquality;    // post-solve quality buffers (pb_hip_quality)
    DeviceAlloc dcons;       // survey-constraint table of the current solve (weights and thresholds are options)
    pbk::Quality Q;
    int lastSolveStatus = PB_SOLVE_NOT_RUN, lastUsable = 0;
    bool batchMode = false;  // the context holds a batch (pb_hip_solve_batch); single-problem entries refuse it
    BatchHost bh;
    FactorTuning tune;
    // Captured PCG iterations per work precision ([0] FP32, [1] FP64): kPcgGraphLaunches launches that take their
    // absolute iteration from the device counter, then the counter's advance. Kept while the launch arguments
    // (problem and work views, stop parameters) stay byte for byte the same; recaptured otherwise.
    struct PcgGraph
    {
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;
    } pcgGraph[2];
    bool pcgGraphs = true;  // options.pcg_graphs of the current solve
    pb_progress_fn progressFn = nullptr;  // SolverControlOptions::progressCallback (pb_hip_set_progress)
    void* progressUser = nullptr;
    int64_t patternBuilds = 0;  // reduced-system patterns built on this context
    bool work32 = false, work64 = false;
    pbk::Problem P;
    pbk::Work<float> W32;
    pbk::Work<double> W64;
    // camera rig: the rig system's structure (a Problem value whose reduced-system fields are the rig's) and its
    // work vectors per precision; the reduced-solve routes run on these
    bool rigSys = false;
    pbk::Problem Psys;
    pbk::Work<float> Wsys32;
    pbk::Work<double> Wsys64;
    std::vector<double> outCapR, outCapC, outSenR, outSenC;
    // Cluster block-Jacobi PCG (pb_cluster.hip): the plan for the current structure and cluster size, used by the
    // PCG route whenever a preconditioner block spans more than one unit.
    struct Cluster
    {
        bool active = false;  // the current solve takes the cluster PCG
        bool wide = false;    // the plan has a cluster of more than one unit
        int size = 0;
        int64_t planBuild = -1;
        DeviceAlloc mem;
        pbk::ClusterDev dev;
    } cluster;
    int costBlocks = 0;
    double* pinned = nullptr;  // host-pinned readback area
    // last results
    std::vector<double> outR, outC, outGroupParams, outPointX;
    bool haveResult = false;
    pb_solve_options lastOptions{};  // options of the last solve (the adjoint linearises with them)
    // Dense reduced solve: matrices of order 64 * ntCap per precision (allocated on first use and when a larger
    // system comes), refinement vectors and the solver's state words and tile flags.
    struct Dense
    {
        DeviceAlloc vec, mat32, mat64;
        int ntCap = 0;
        float* A32 = nullptr;
        double* A64 = nullptr;
        float* v32 = nullptr;
        double* v64 = nullptr;
        double *x = nullptr, *r = nullptr, *rowNorm = nullptr, *norm = nullptr;
        pbk::DenseState st;
    } dense;
    // Block-tridiagonal reduced solve (pb_band.hip): the plan of the current reduced-system pattern (made on
    // first use after a pattern build) and its arenas per precision.
    struct Band
    {
        int64_t planBuild = -1;  // patternBuilds value the plan belongs to
        bool ok = false;         // the pattern fits the band route
        bool active = false;     // the current solve (or adjoint) takes the band route
        std::string why;         // why the pattern does not fit, when it does not
        int bandwidth = 0;       // pose units, in the bandwidth-reducing order
        pbk::BandView view;
        DeviceAlloc tables, arena32, arena64;
        float* Z32 = nullptr;
        double* Z64 = nullptr;
    } band;
    // profiling
    bool profile = false;
    int64_t pcgIterationsProfiled = 0;
    double lastScalarBytes = 8.0;
    std::array<Family, kFamilyCount> fam;
    std::vector<hipEvent_t> eventPool;
    size_t eventNext = 0;
    hipEvent_t trialEv[2] = {nullptr, nullptr};  // brackets the trial-state kernels of a line-search pass
    hipEvent_t stepEv[7] = {};  // phase boundaries of a step enqueued with one host wait (fusedStep)
};

namespace
{
    // elapsed device time between two recorded events, once the later one has completed
    double eventSeconds(hipEvent_t a, hipEvent_t b)
    {
        float ms = 0.0f;
        return hipEventElapsedTime(&ms, a, b) == hipSuccess ? 1e-3 * ms : 0.0;
    }

    hipEvent_t takeEvent(pb_ctx* ctx)
    {
        if (ctx->eventNext == ctx->eventPool.size())
        {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            ctx->eventPool.push_back(e);
        }
        return ctx->eventPool[ctx->eventNext++];
    }

    struct Span
    {
        pb_ctx* ctx;
        int family;
        hipEvent_t stop = nullptr;
        Span(pb_ctx* c, int f, bool sample = true) : ctx(c), family(f)
        {
            if (!ctx->profile || !sample) return;
            hipEvent_t start = takeEvent(ctx);
            stop = takeEvent(ctx);
            check(hipEventRecord(start, ctx->stream), "hipEventRecord");
            ctx->fam[family].pending.push_back({start, stop, 1});
        }
        ~Span()
        {
            if (ctx->profile && stop) (void)hipEventRecord(stop, ctx->stream);
        }
    };

    void collectProfile(pb_ctx* ctx)
    {
        if (!ctx->profile) return;
        check(hipStreamSynchronize(ctx->stream), "sync");
        for (Family& f : ctx->fam)
        {
            for (const TimedSpan& sp : f.pending)
            {
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, sp.start, sp.stop), "hipEventElapsedTime");
                f.ms += ms;
                f.launches += sp.launches;
            }
            f.pending.clear();
        }
        ctx->eventNext = 0;
    }

    // Cholesky factor (lower, row-major) of a symmetric 3x3 matrix; false unless every pivot is positive.
    bool cholesky3(const double* m, double* L)
    {
        for (int i = 0; i < 9; ++i)
        {
            if (!std::isfinite(m[i])) return false;
            L[i] = 0.0;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j)
            {
                double v = 0.5 * (m[3 * i + j] + m[3 * j + i]);
                for (int k = 0; k < j; ++k) v -= L[3 * i + k] * L[3 * j + k];
                if (i == j)
                {
                    if (!(v > 0.0)) return false;
                    L[3 * i + i] = std::sqrt(v);
                }
                else
                {
                    L[3 * i + j] = v / L[3 * j + j];
                }
            }
        return true;
    }

    // Whitening matrix of a control point (makeControlPointWhitening, src/control_point.cpp:91-165): residual
    // A (X - P) with A = diag(1/sigma), the inverse Cholesky factor of a covariance, or a square-root
    // information matrix whose A^T A is positive definite.
    bool controlWhitening(int kind, double sigma, const double* m, double* A)
    {
        for (int i = 0; i < 9; ++i) A[i] = 0.0;
        if (kind == PB_CONTROL_ISOTROPIC_SIGMA)
        {
            if (!std::isfinite(sigma) || sigma <= 0.0) return false;
            A[0] = A[4] = A[8] = 1.0 / sigma;
            return true;
        }
        if (!m) return false;
        double L[9];
        if (kind == PB_CONTROL_COVARIANCE)
        {
            if (!cholesky3(m, L)) return false;
            // forward substitution L A = I, column by column
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                {
                    double v = r == c ? 1.0 : 0.0;
                    for (int k = 0; k < r; ++k) v -= L[3 * r + k] * A[3 * k + c];
                    A[3 * r + c] = v / L[3 * r + r];
                }
            return true;
        }
        if (kind == PB_CONTROL_SQRT_INFORMATION)
        {
            double I[9] = {};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) I[3 * i + j] += m[3 * k + i] * m[3 * k + j];
            if (!cholesky3(I, L)) return false;
            std::copy(m, m + 9, A);
            return true;
        }
        return false;
    }

    // The survey constraints of a view, validated as validateProblem does (src/problem.cpp:241-259).
    std::string copyConstraints(const pb_problem_view* v, Input* r)
    {
        auto finite3 = [](const double* q) { return std::isfinite(q[0]) && std::isfinite(q[1]) && std::isfinite(q[2]); };
        for (int i = 0; i < v->num_control_points; ++i)
        {
            Input::Constraint k;
            k.track = v->cp_track[i];
            k.index = i;
            std::copy(v->cp_point + 3 * i, v->cp_point + 3 * i + 3, k.target);
            k.weight = v->cp_weight ? v->cp_weight[i] : 1.0;
            if (k.track < 0 || k.track >= v->num_tracks) return "control point " + std::to_string(i) + ": track out of range";
            const int kind = v->cp_uncertainty ? v->cp_uncertainty[i] : PB_CONTROL_ISOTROPIC_SIGMA;
            if (!finite3(k.target) || !std::isfinite(k.weight) || k.weight <= 0.0 ||
                !controlWhitening(kind, v->cp_sigma ? v->cp_sigma[i] : 1.0, v->cp_matrix ? v->cp_matrix + 9 * i : nullptr,
                                  k.A))
                return "track contains an invalid control-point constraint (control point " + std::to_string(i) + ")";
            r->cons.push_back(k);
        }
        for (int i = 0; i < v->num_laser_planes; ++i)
        {
            Input::Constraint k;
            k.track = v->lp_track[i];
            k.plane = true;
            k.index = i;
            std::copy(v->lp_point + 3 * i, v->lp_point + 3 * i + 3, k.target);
            std::copy(v->lp_normal + 3 * i, v->lp_normal + 3 * i + 3, k.A);
            k.weight = v->lp_weight ? v->lp_weight[i] : 1.0;
            if (k.track < 0 || k.track >= v->num_tracks) return "laser plane " + std::to_string(i) + ": track out of range";
            const double nn = std::hypot(k.A[0], k.A[1], k.A[2]);
            if (!finite3(k.target) || !finite3(k.A) || !std::isfinite(nn) || nn <= 0.0 || !std::isfinite(k.weight) ||
                k.weight <= 0.0)
                return "track contains an invalid laser-plane constraint (laser plane " + std::to_string(i) + ")";
            r->cons.push_back(k);
        }
        r->nControl = v->num_control_points;
        r->nPlanes = v->num_laser_planes;
        return {};
    }

    // A rig's poses and bindings (validated by validateInput). As runConcreteBackend and makeSolverOptions do
    // before the solve: the bound cameras are composed from their capture and sensor (solver.cpp:282), each sensor
    // is a calibration group when no group ids are given and unbound cameras take group 0 (Conversion.cpp:83-98),
    // and a camera whose capture and sensor are both fixed joins the fixed cameras (Conversion.cpp:120-142).
    void copyRig(const pb_problem_view* v, Input* r)
    {
        r->nCap = v->num_rig_captures;
        r->nSen = v->num_rig_sensors;
        r->capR.assign(v->rig_capture_rotation, v->rig_capture_rotation + 9 * size_t(r->nCap));
        r->capC.assign(v->rig_capture_center, v->rig_capture_center + 3 * size_t(r->nCap));
        r->senR.assign(v->rig_sensor_rotation, v->rig_sensor_rotation + 9 * size_t(r->nSen));
        r->senC.assign(v->rig_sensor_center, v->rig_sensor_center + 3 * size_t(r->nSen));
        r->capFixed.assign(r->nCap, 0);
        r->senFixed.assign(r->nSen, 0);
        for (int k = 0; k < r->nCap; ++k) r->capFixed[k] = v->rig_capture_fixed && v->rig_capture_fixed[k] ? 1 : 0;
        for (int k = 0; k < r->nSen; ++k) r->senFixed[k] = v->rig_sensor_fixed && v->rig_sensor_fixed[k] ? 1 : 0;
        r->camCapture.assign(v->cam_rig_capture, v->cam_rig_capture + r->nCam);
        r->camSensor.assign(v->cam_rig_sensor, v->cam_rig_sensor + r->nCam);
        for (int c = 0; c < r->nCam; ++c)
            pbdev::composeRig(&r->capR[9 * r->camCapture[c]], &r->capC[3 * r->camCapture[c]],
                              &r->senR[9 * r->camSensor[c]], &r->senC[3 * r->camSensor[c]], &r->R[9 * c], &r->C[3 * c]);
        if (!r->hasGroup)
        {
            r->hasGroup = true;
            r->group.assign(r->nCam, 0);
            std::vector<int32_t> keys;
            for (int c = 0; c < r->nCam; ++c)
            {
                auto it = std::find(keys.begin(), keys.end(), r->camSensor[c]);
                if (it == keys.end()) it = keys.insert(keys.end(), r->camSensor[c]);
                r->group[c] = static_cast<int32_t>(it - keys.begin());
            }
        }
        for (int c = 0; c < r->nCam; ++c)
            if (r->capFixed[r->camCapture[c]] && r->senFixed[r->camSensor[c]] &&
                std::find(r->fixedCams.begin(), r->fixedCams.end(), c) == r->fixedCams.end())
                r->fixedCams.push_back(c);
    }

    // A rigid rotation: finite, orthonormal to 1e-6 and proper.
    bool rotationValid(const double* R)
    {
        for (int i = 0; i < 9; ++i)
            if (!std::isfinite(R[i])) return false;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                const double d = R[i] * R[j] + R[3 + i] * R[3 + j] + R[6 + i] * R[6 + j];
                if (std::abs(d - (i == j ? 1.0 : 0.0)) > 1e-6) return false;
            }
        return R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                   R[2] * (R[3] * R[7] - R[4] * R[6]) >
               0.0;
    }

    // The rig must compose: every camera bound to a capture and a sensor in range (the reference indexes its
    // block tables with the binding, BundleAdjustPlaMatrixProblem.cpp:192-198), and every pose a rigid transform.
    std::string validateRig(const pb_problem_view* v)
    {
        if (v->num_rig_captures < 0 || v->num_rig_sensors < 0) return "negative rig counts";
        if (v->num_rig_captures == 0 && v->num_rig_sensors == 0) return {};
        if (!v->rig_capture_rotation || !v->rig_capture_center || !v->rig_sensor_rotation || !v->rig_sensor_center ||
            !v->cam_rig_capture || !v->cam_rig_sensor)
            return "rig arrays are NULL";
        auto finite3 = [](const double* q) { return std::isfinite(q[0]) && std::isfinite(q[1]) && std::isfinite(q[2]); };
        for (int k = 0; k < v->num_rig_captures; ++k)
            if (!rotationValid(v->rig_capture_rotation + 9 * k) || !finite3(v->rig_capture_center + 3 * k))
                return "rig capture " + std::to_string(k) + " is not a valid rigid transform";
        for (int k = 0; k < v->num_rig_sensors; ++k)
            if (!rotationValid(v->rig_sensor_rotation + 9 * k) || !finite3(v->rig_sensor_center + 3 * k))
                return "rig sensor " + std::to_string(k) + " is not a valid rigid transform";
        for (int c = 0; c < v->num_cameras; ++c)
        {
            const int k = v->cam_rig_capture[c], q = v->cam_rig_sensor[c];
            if (k < 0 || k >= v->num_rig_captures || q < 0 || q >= v->num_rig_sensors)
                return "rig binding of camera " + std::to_string(c) + " is missing or out of range";
        }
        return {};
    }

    void copyIn(const pb_problem_view* v, Input* in)
    {
        Input r;
        r.nCam = v->num_cameras;
        r.nTracks = v->num_tracks;
        const size_t n = static_cast<size_t>(r.nCam);
        r.R.assign(v->cam_rotation, v->cam_rotation + 9 * n);
        r.C.assign(v->cam_center, v->cam_center + 3 * n);
        r.slots.assign(v->cam_intrinsics, v->cam_intrinsics + 13 * n);
        r.ref = v->ref_intrinsics ? std::vector<double>(v->ref_intrinsics, v->ref_intrinsics + 13 * n) : r.slots;
        r.conv = v->cam_convention ? std::vector<int32_t>(v->cam_convention, v->cam_convention + n)
                                   : std::vector<int32_t>(n, PB_BROWN_OPENCV);
        r.hasGroup = v->cam_group != nullptr;
        if (r.hasGroup) r.group.assign(v->cam_group, v->cam_group + n);
        if (v->cam_image_size) r.imageSize.assign(v->cam_image_size, v->cam_image_size + 2 * n);
        r.X.assign(v->track_point, v->track_point + 3 * static_cast<size_t>(r.nTracks));
        r.off.assign(v->track_offset, v->track_offset + r.nTracks + 1);
        const int64_t no = r.off.back();
        r.obsCam.assign(v->obs_camera, v->obs_camera + no);
        r.uv.assign(v->obs_uv, v->obs_uv + 2 * no);
        r.weight = v->obs_weight ? std::vector<double>(v->obs_weight, v->obs_weight + no) : std::vector<double>(no, 1.0);
        r.scale = v->obs_scale ? std::vector<double>(v->obs_scale, v->obs_scale + no) : std::vector<double>(no, 1.0);
        if (v->fixed_cameras) r.fixedCams.assign(v->fixed_cameras, v->fixed_cameras + v->num_fixed_cameras);
        if (v->fixed_tracks) r.fixedTracks.assign(v->fixed_tracks, v->fixed_tracks + v->num_fixed_tracks);
        r.gaugePolicy = v->gauge_policy;
        r.gaugeAnchor = v->gauge_anchor_camera;
        r.gaugeScale = v->gauge_scale_camera;
        r.gaugeBaseline = v->gauge_baseline;
        const std::string err = copyConstraints(v, &r);
        if (!err.empty()) throw std::invalid_argument(err);
        if (v->num_rig_captures > 0 || v->num_rig_sensors > 0) copyRig(v, &r);
        *in = std::move(r);
    }

    std::string validateInput(const pb_problem_view* v)
    {
        if (!v) return "problem is NULL";
        if (v->num_cameras < 0 || v->num_tracks < 0) return "negative counts";
        if (v->num_cameras > 0 && (!v->cam_rotation || !v->cam_center || !v->cam_intrinsics))
            return "camera arrays are NULL";
        if (v->num_tracks > 0 && (!v->track_point || !v->track_offset || !v->obs_camera || !v->obs_uv))
            return "track arrays are NULL";
        if (v->num_tracks > 0)
        {
            if (v->track_offset[0] != 0) return "track_offset[0] must be 0";
            for (int t = 0; t < v->num_tracks; ++t)
                if (v->track_offset[t + 1] < v->track_offset[t]) return "track_offset must be non-decreasing";
        }
        const std::string rerr = validateRig(v);
        if (!rerr.empty()) return rerr;
        const bool rig = v->num_rig_captures > 0 || v->num_rig_sensors > 0;
        for (int c = 0; c < v->num_cameras; ++c)
        {
            // a rig camera's pose comes from its capture and sensor
            for (int k = 0; k < 9 && !rig; ++k)
                if (!std::isfinite(v->cam_rotation[9 * c + k])) return "camera pose must be finite";
            for (int k = 0; k < 3 && !rig; ++k)
                if (!std::isfinite(v->cam_center[3 * c + k])) return "camera pose must be finite";
            const double fx = v->cam_intrinsics[13 * c];
            if (!std::isfinite(fx) || fx <= 0.0 || !std::isfinite(v->cam_intrinsics[13 * c + 1]))
                return "camera intrinsics must be finite and positive";
            if (v->cam_convention && v->cam_convention[c] != PB_BROWN_OPENCV && v->cam_convention[c] != PB_BROWN_METASHAPE)
                return "camera convention must be OpenCV or Metashape";
            if (v->cam_group && v->cam_group[c] < 0) return "calibration group ids must be non-negative";
        }
        if (v->num_fixed_cameras > 0 && !v->fixed_cameras) return "fixed_cameras is NULL";
        if (v->num_fixed_tracks > 0 && !v->fixed_tracks) return "fixed_tracks is NULL";
        if (v->gauge_policy < 0 || v->gauge_policy > 2) return "gauge policy is invalid";
        if (v->num_control_points < 0 || v->num_laser_planes < 0) return "negative constraint counts";
        if (v->num_control_points > 0 && (!v->cp_track || !v->cp_point)) return "control-point arrays are NULL";
        if (v->num_control_points > 0 && v->cp_uncertainty && !v->cp_matrix)
            for (int i = 0; i < v->num_control_points; ++i)
                if (v->cp_uncertainty[i] != PB_CONTROL_ISOTROPIC_SIGMA) return "cp_matrix is NULL";
        if (v->num_laser_planes > 0 && (!v->lp_track || !v->lp_point || !v->lp_normal)) return "laser-plane arrays are NULL";
        return {};
    }

    std::string validateOptions(const pb_solve_options* o)
    {
        if (!o) return "options are NULL";
        if (o->max_iterations <= 0) return "max_iterations must be positive";
        if (o->robust_loss < 0 || o->robust_loss > 2) return "robust loss is invalid";
        if (!std::isfinite(o->robust_loss_scale) || o->robust_loss_scale <= 0.0) return "robust loss scale must be positive";
        if (!(o->armijo_coefficient > 0.0) || !(o->armijo_coefficient < 1.0) || o->line_search_steps <= 0 ||
            o->line_search_steps > 64)
            return "Armijo parameters must satisfy coefficient in (0,1) and steps in [1,64]";
        if (!std::isfinite(o->max_initial_track_rms) || o->max_initial_track_rms < 0.0)
            return "max_initial_track_rms must be finite and non-negative";
        if (o->shared_intrinsic_mask & ~0x1FFFu) return "shared_intrinsic_mask has bits beyond slot 12";
        if ((o->shared_intrinsic_mask & 0x1FFEu) && !(o->shared_intrinsic_mask & 1u) && o->shared_intrinsic_mask != (1u << 4))
            return "refining aspect, principal point or distortion requires refining the shared focal length";
        if (o->precision < PB_PRECISION_FP64 || o->precision > PB_PRECISION_FP32_SEED) return "precision is invalid";
        if (o->reduced_solver < PB_REDUCED_AUTO || o->reduced_solver > PB_REDUCED_BAND) return "reduced_solver is invalid";
        if (o->dense_max_rows < 0) return "dense_max_rows must be non-negative";
        // plaMatrixPreconditionerClusterSize must be in [1, 16] (src/options.cpp:393-394)
        if (o->preconditioner_cluster_size < 1 || o->preconditioner_cluster_size > 16)
            return "plaMatrixPreconditionerClusterSize must be in [1, 16]";
        if (o->pcg_relative_tolerance < 0.0 || o->pcg_absolute_tolerance < 0.0 || o->pcg_max_iterations < 0)
            return "PCG controls must be non-negative";
        if (o->batch_factor_tiles != 0 && o->batch_factor_tiles != 1) return "batch_factor_tiles must be 0 or 1";
        if (o->batch_factor_shape < -1 || o->batch_factor_shape >= pbk::kFactorShapes)
            return "batch_factor_shape must be -1 (timed) or a shape index below " + std::to_string(pbk::kFactorShapes);
        return {};
    }

    void releaseStructure(pb_ctx* ctx);
    void uploadStructure(pb_ctx* ctx, Prepared&& p, Layout&& L, const pb_solve_options& o);

    // The options a prepared structure depends on: the active problem (refined pose and slots, initial-RMS
    // gate) and the intrinsic groups (bounds, prior sigmas, transition prior).
    bool sameStructureOptions(const pb_solve_options& a, const pb_solve_options& b)
    {
        bool same = a.shared_intrinsic_mask == b.shared_intrinsic_mask && a.refine_camera_pose == b.refine_camera_pose &&
                    a.max_initial_track_rms == b.max_initial_track_rms && a.min_focal_scale == b.min_focal_scale &&
                    a.max_focal_scale == b.max_focal_scale && a.min_aspect_scale == b.min_aspect_scale &&
                    a.max_aspect_scale == b.max_aspect_scale &&
                    a.max_principal_offset_fraction == b.max_principal_offset_fraction &&
                    a.focal_prior_sigma == b.focal_prior_sigma &&
                    a.principal_prior_sigma_fraction == b.principal_prior_sigma_fraction &&
                    a.aspect_prior_sigma == b.aspect_prior_sigma &&
                    a.low_order_distortion_scale == b.low_order_distortion_scale &&
                    a.use_transition_prior == b.use_transition_prior &&
                    a.previous_intrinsic_mask == b.previous_intrinsic_mask;
        for (int i = 0; i < 9; ++i) same = same && a.max_abs[i] == b.max_abs[i] && a.prior_sigma[i] == b.prior_sigma[i];
        return same;
    }

    // Builds the prepared structure and uploads it (setup; once per problem and structural options).
    void ensureStructure(pb_ctx* ctx, const pb_solve_options& o)
    {
        if (ctx->structureValid && sameStructureOptions(ctx->structureOptions, o))
            return;
        releaseStructure(ctx);
        Prepared p;
        const std::string gerr = normalizeGauge(ctx->in, o, &p);
        if (!gerr.empty()) throw std::invalid_argument(gerr);
        prepareActive(ctx->in, o, &p);
        prepareGroups(ctx->in, o, &p);
        Layout L;
        buildLayout(ctx->in, p, &L);
        if (p.rig) buildSysLayout(ctx->in, p, &L);
        uploadStructure(ctx, std::move(p), std::move(L), o);
    }

    void releaseStructure(pb_ctx* ctx)
    {
        ctx->structureValid = false;
        ctx->dstruct.release();
        ctx->dwork32.release();
        ctx->dwork64.release();
        ctx->dbatch.release();
        ctx->dquality.release();
        ctx->dcons.release();
        ctx->Q = pbk::Quality{};
        ctx->work32 = ctx->work64 = false;
        ctx->cluster.mem.release();
        ctx->cluster = pb_ctx::Cluster{};
    }

    // Device copy of a prepared problem and its layout (setup).
    // The fused-PCG wave slots of a reduced system's rows (slot table and the first column ids of each slot).
    void uploadCgSlots(DeviceAlloc& A, const Rows& R, pbk::Problem* P, hipStream_t s)
    {
        P->cgBlocks = static_cast<int>(R.cgWaveRow.size()) / pbk::kCgWaves;
        std::vector<int32_t> slots(4 * R.cgWaveRow.size());
        for (size_t i = 0; i < R.cgWaveRow.size(); ++i)
        {
            slots[4 * i] = R.cgWaveRow[i];
            slots[4 * i + 1] = R.cgWaveBegin[i];
            slots[4 * i + 2] = R.cgWaveEnd[i];
            slots[4 * i + 3] = R.cgWaveLead[i];
        }
        P->cgSlot = reinterpret_cast<const int4*>(A.upload(slots, s));
        std::vector<int32_t> direct(pbk::kCgDirect * R.cgWaveRow.size(), -1);
        for (size_t i = 0; i < R.cgWaveRow.size(); ++i)
            for (int j = 0; j < pbk::kCgDirect && R.cgWaveBegin[i] + j < R.cgWaveEnd[i]; ++j)
                direct[pbk::kCgDirect * i + j] = R.colIdx[R.cgWaveBegin[i] + j];
        P->cgColDirect = A.upload(direct, s);
    }

    void uploadStructure(pb_ctx* ctx, Prepared&& p, Layout&& L, const pb_solve_options& o)
    {
        ++ctx->patternBuilds;
        hipStream_t s = ctx->stream;
        DeviceAlloc& A = ctx->dstruct;
        pbk::Problem P;
        P.nCam = ctx->in.nCam;
        P.nGroup = p.G;
        P.nObs = L.nObs;
        P.nActive = L.nActive;
        P.nElim = L.nElim;
        P.nEntries = L.nEntries;
        P.nUnits = p.nUnits;
        P.nBlocks = L.nBlocks;
        P.nNnz = L.nNnz;
        P.nItems = L.nItems;
        P.nPairs = L.nPairs;
        P.maxGroupK = L.maxGroupK;
        P.gaugeAnchor = p.anchor;
        P.gaugeScale = (p.anchor >= 0 && p.scaleCam >= 0 && p.baseline > 0.0) ? p.scaleCam : -1;
        P.gaugeBaseline = p.baseline;
        P.globalDim = p.globalDim;
        P.camR = A.upload(ctx->in.R, s);
        P.camC = A.upload(ctx->in.C, s);
        P.camSlots = A.upload(ctx->in.slots, s);
        P.refSlots = A.upload(ctx->in.ref, s);
        P.camConv = A.upload(ctx->in.conv, s);
        P.camGroup = A.upload(L.camGroup, s);
        P.camUnit = A.upload(L.camUnit, s);
        P.cams = A.alloc<pbdev::Camera>(P.nCam);
        P.camObsOffset = A.upload(L.camObsOffset, s);
        P.camObs = A.upload(L.camObs, s);
        {
            std::vector<int32_t> cp(L.nObs);
            std::vector<double> cuv(size_t(L.nObs) * 2), cw(L.nObs);
            for (int k = 0; k < L.nObs; ++k)
            {
                const int o = L.camObs[k];
                cp[k] = L.obsPoint[o];
                cuv[2 * size_t(k)] = L.obsUV[2 * size_t(o)];
                cuv[2 * size_t(k) + 1] = L.obsUV[2 * size_t(o) + 1];
                cw[k] = L.obsW[o];
            }
            P.camObsPoint = A.upload(cp, s);
            P.camObsUV = A.upload(cuv, s);
            P.camObsW = A.upload(cw, s);
        }
        P.camDiagBlock = A.upload(L.camDiagBlock, s);
        P.camGroupBlock = A.upload(L.camGroupBlock, s);
        P.groupParams = A.upload(p.gParams, s);
        P.groupPrior = A.upload(p.gPrior, s);
        P.groupInvSigma = A.upload(p.gInvSigma, s);
        P.groupLower = A.upload(p.gLower, s);
        P.groupUpper = A.upload(p.gUpper, s);
        P.groupEnabled = A.upload(p.gEnabled, s);
        P.groupActive = A.upload(p.gActive, s);
        P.groupK = A.upload(p.gK, s);
        P.groupUnit = A.upload(p.gUnit, s);
        P.groupCamOffset = A.upload(L.groupCamOffset, s);
        P.groupCams = A.upload(L.groupCams, s);
        P.groupPairBlock = A.upload(L.groupPairBlock, s);
        P.slotToCompact = A.upload(p.slotToCompact, s);
        P.obsCam = A.upload(L.obsCam, s);
        P.obsPoint = A.upload(L.obsPoint, s);
        P.obsUV = A.upload(L.obsUV, s);
        P.obsW = A.upload(L.obsW, s);
        P.obsPoseEntry = A.upload(L.obsPoseEntry, s);
        P.obsGroupEntry = A.upload(L.obsGroupEntry, s);
        P.pointX = A.upload(L.pointX, s);
        P.obsOffset = A.upload(L.obsOffset, s);
        P.pointElim = A.upload(L.pointElim, s);
        P.elimTrack = A.upload(L.elimTrack, s);
        P.entryOffset = A.upload(L.entryOffset, s);
        P.entryUnit = A.upload(L.entryUnit, s);
        P.entryPoint = A.upload(L.entryPoint, s);
        P.nLinWaves = static_cast<int>(L.linLaneObs.size() / pbk::kWaveLanes);
        P.nLinFallback = static_cast<int>(L.linFallback.size());
        P.linLaneObs = A.upload(L.linLaneObs, s);
        P.linLanePoint = A.upload(L.linLanePoint, s);
        P.linFallback = A.upload(L.linFallback, s);
        P.blockRow = A.upload(L.blockRow, s);
        P.blockCol = A.upload(L.blockCol, s);
        P.blockPosUp = A.upload(L.rows.blockPosUp, s);
        P.blockPosLo = A.upload(L.rows.blockPosLo, s);
        P.blockItemOffset = A.upload(L.blockItemOffset, s);
        P.itemBlock = A.upload(L.itemBlock, s);
        P.itemBegin = A.upload(L.itemBegin, s);
        P.pairE1 = A.upload(L.pairE1, s);
        P.pairE2 = A.upload(L.pairE2, s);
        P.rowPtr = A.upload(L.rows.rowPtr, s);
        P.colIdx = A.upload(L.rows.colIdx, s);
        P.diagPos = A.upload(L.rows.diagPos, s);
        P.unitWidth = A.upload(p.unitWidth, s);
        {
            // post-solve quality: input cameras (own intrinsics), input points, strict-RMS exclusions
            std::vector<pbdev::Camera> inCams(std::max(1, ctx->in.nCam));
            for (int c = 0; c < ctx->in.nCam; ++c) inCams[c] = inputCamera(ctx->in, c);
            P.inputCams = A.upload(inCams, s);
            P.pointX0 = A.upload(L.pointX, s);
            std::vector<int8_t> bad(std::max(1, L.nActive), 0);
            for (int a = 0; a < L.nActive; ++a)
            {
                const int t = L.activeTrackOf[a];
                for (int64_t i = ctx->in.off[t]; i < ctx->in.off[t + 1]; ++i)
                {
                    const double sc = sanitizedScale(ctx->in.scale[i]);
                    const bool weighted = sanitizedWeight(ctx->in.weight[i]) / (sc * sc) > 0.0;
                    if (weighted && !usable(ctx->in, i)) bad[a] = 1;
                }
            }
            P.trackStrictBad = A.upload(bad, s);
        }
        uploadCgSlots(A, L.rows, &P, s);
        check(hipStreamSynchronize(s), "structure upload");
        L.pairE1.clear();
        L.pairE1.shrink_to_fit();
        L.pairE2.clear();
        L.pairE2.shrink_to_fit();
        L.linLaneObs = std::vector<int32_t>();
        L.linLanePoint = std::vector<int32_t>();
        ctx->rigSys = p.rig;
        if (p.rig)
        {
            const Input& in = ctx->in;
            P.nCap = in.nCap;
            P.nSen = in.nSen;
            P.capR = A.upload(in.capR, s);
            P.capC = A.upload(in.capC, s);
            P.senR = A.upload(in.senR, s);
            P.senC = A.upload(in.senC, s);
            P.camCapture = A.upload(in.camCapture, s);
            P.camSensor = A.upload(in.camSensor, s);
            P.capUnit = A.upload(p.rigCapBlock, s);  // rig units are the reference's rig block ids
            P.senUnit = A.upload(p.rigSenBlock, s);
            P.rigMap = A.alloc<double>(72 * size_t(std::max(1, P.nCam)));
            P.nSysUnits = L.nSysUnits;
            P.nSysBlocks = static_cast<int>(L.sysBlockRow.size());
            P.sysBlockRow = A.upload(L.sysBlockRow, s);
            P.sysBlockCol = A.upload(L.sysBlockCol, s);
            P.sysPosUp = A.upload(L.sysRows.blockPosUp, s);
            P.sysPosLo = A.upload(L.sysRows.blockPosLo, s);
            P.projOffset = A.upload(L.projOffset, s);
            P.projPos = A.upload(L.projPos, s);
            P.projMapL = A.upload(L.projMapL, s);
            P.projMapR = A.upload(L.projMapR, s);
            P.sysRhsOffset = A.upload(L.sysRhsOffset, s);
            P.sysRhsUnit = A.upload(L.sysRhsUnit, s);
            P.sysRhsMap = A.upload(L.sysRhsMap, s);
            P.expUnit = A.upload(L.expUnit, s);
            P.expMap = A.upload(L.expMap, s);
            // a rig has no gauge tangent: the scale camera's capture is held (Assembly.cpp:204)
            P.gaugeScale = -1;
            pbk::Problem Ps = P;
            Ps.nUnits = L.nSysUnits;
            Ps.nBlocks = P.nSysBlocks;
            Ps.nNnz = L.sysRows.nNnz;
            Ps.blockRow = P.sysBlockRow;
            Ps.blockCol = P.sysBlockCol;
            Ps.blockPosUp = P.sysPosUp;
            Ps.blockPosLo = P.sysPosLo;
            Ps.rowPtr = A.upload(L.sysRows.rowPtr, s);
            Ps.colIdx = A.upload(L.sysRows.colIdx, s);
            Ps.diagPos = A.upload(L.sysRows.diagPos, s);
            Ps.unitWidth = A.upload(L.sysUnitWidth, s);
            Ps.globalDim = L.sysGlobalDim;
            uploadCgSlots(A, L.sysRows, &Ps, s);
            check(hipStreamSynchronize(s), "rig structure upload");
            ctx->Psys = Ps;
        }
        ctx->P = P;
        ctx->prep = std::move(p);
        ctx->lay = std::move(L);
        ctx->structureValid = true;
        ctx->structureOptions = o;
        ctx->costBlocks = std::max(1, std::min(2048, (P.nObs + 255) / 256));
    }

    template <typename T>
    void allocWork(pb_ctx* ctx, DeviceAlloc& A, pbk::Work<T>& W)
    {
        const pbk::Problem& P = ctx->P;
        const int nbElim = std::max(1, (P.nElim + 255) / 256);
        const int nParts = nbElim;  // back-substitution block partials
        W.W = A.alloc<T>(size_t(P.nEntries) * 18);
        W.Hinv = A.alloc<T>(size_t(P.nElim) * 9);
        W.bp = A.alloc<double>(size_t(P.nElim) * 3);
        W.y = A.alloc<double>(size_t(P.nElim) * 3);
        W.dp = A.alloc<double>(size_t(P.nElim) * 3);
        // sized as the builds clear them (at least one block / unit, also for problems with nothing to refine)
        W.D = A.alloc<double>(size_t(std::max(1, P.nBlocks)) * 36);
        W.directRhs = A.alloc<double>(size_t(std::max(1, P.nUnits)) * 6);
        W.rhs = A.alloc<double>(size_t(P.nUnits) * 6);
        W.camCost = A.alloc<double>(P.nCam);
        W.camGroupPartial = A.alloc<double>(size_t(P.nCam) * pbk::kGroupPartialStride);
        W.groupPriorCost = A.alloc<double>(std::max(1, P.nGroup));
        W.itemPartial = A.alloc<T>(size_t(P.nItems) * 36);
        W.itemPartialRhs = A.alloc<double>(size_t(P.nItems) * 6);
        W.S = A.alloc<T>(size_t(P.nNnz) * 36);
        W.Minv = A.alloc<T>(size_t(P.nUnits) * 36);
        W.x = A.alloc<T>(size_t(P.nUnits) * 6);
        W.p = A.alloc<T>(size_t(P.nUnits) * 6);
        for (int k = 0; k < 2; ++k)
        {
            W.cgR[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgU[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgW[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgMW[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgS[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgMS[k] = A.alloc<T>(size_t(P.nUnits) * 6);
            W.cgPart[k] = A.alloc<double>(size_t(P.nUnits) * 3);
        }
        W.partDd = A.alloc<double>(nParts);
        W.partSs = A.alloc<double>(nParts);
        W.scalars = A.alloc<double>(pbk::kScalarCount);
        W.flags = A.alloc<int32_t>(pbk::kFlagCount);
        W.iflags = A.alloc<int32_t>(pbk::kIflagCount);
        W.candGroupParams = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nGroup) * 13);
        W.candPriorCost = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nGroup));
        W.candR = A.alloc<double>(size_t(pbk::kMaxAlpha) * P.nCam * 9);
        W.candC = A.alloc<double>(size_t(pbk::kMaxAlpha) * P.nCam * 3);
        W.candCams = A.alloc<pbdev::Camera>(size_t(pbk::kMaxAlpha) * P.nCam);
        W.alphas = A.alloc<double>(pbk::kMaxAlpha);
        W.trialPart = A.alloc<double>(size_t(ctx->costBlocks) * pbk::kMaxAlpha);
        W.trialBad = A.alloc<int32_t>(size_t(ctx->costBlocks) * pbk::kMaxAlpha);
        W.trialCost = A.alloc<double>(pbk::kMaxAlpha);
        check(hipMemsetAsync(W.camGroupPartial, 0, sizeof(double) * size_t(P.nCam) * pbk::kGroupPartialStride,
                             ctx->stream),
              "memset");
    }

    // Work of a rig problem's rig system: its own matrix, preconditioner, solution and PCG vectors; scalars,
    // flags and the trial buffers are the camera-level work's. The camera-level work gets the candidate rig poses
    // and the rig solution it reads in the line search.
    template <typename T>
    void allocSysWork(pb_ctx* ctx, DeviceAlloc& A, pbk::Work<T>& W, pbk::Work<T>& Ws)
    {
        const pbk::Problem& P = ctx->P;
        const pbk::Problem& Ps = ctx->Psys;
        const size_t nu = size_t(std::max(1, Ps.nUnits));
        W.candCapR = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nCap) * 9);
        W.candCapC = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nCap) * 3);
        W.candSenR = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nSen) * 9);
        W.candSenC = A.alloc<double>(size_t(pbk::kMaxAlpha) * std::max(1, P.nSen) * 3);
        Ws = W;
        Ws.S = A.alloc<T>(size_t(std::max(1, Ps.nNnz)) * 36);
        Ws.Minv = A.alloc<T>(nu * 36);
        Ws.x = A.alloc<T>(nu * 6);
        Ws.p = A.alloc<T>(nu * 6);
        Ws.rhs = A.alloc<double>(nu * 6);
        Ws.directRhs = nullptr;
        for (int k = 0; k < 2; ++k)
        {
            Ws.cgR[k] = A.alloc<T>(nu * 6);
            Ws.cgU[k] = A.alloc<T>(nu * 6);
            Ws.cgW[k] = A.alloc<T>(nu * 6);
            Ws.cgMW[k] = A.alloc<T>(nu * 6);
            Ws.cgS[k] = A.alloc<T>(nu * 6);
            Ws.cgMS[k] = A.alloc<T>(nu * 6);
            Ws.cgPart[k] = A.alloc<double>(nu * 3);
        }
        W.rigX = Ws.x;
        Ws.rigX = Ws.x;
    }

    template <typename T>
    pbk::Work<T>& work(pb_ctx* ctx);
    template <>
    pbk::Work<float>& work<float>(pb_ctx* ctx)
    {
        if (!ctx->work32)
        {
            allocWork(ctx, ctx->dwork32, ctx->W32);
            if (ctx->rigSys) allocSysWork(ctx, ctx->dwork32, ctx->W32, ctx->Wsys32);
            ctx->work32 = true;
        }
        return ctx->W32;
    }
    template <>
    pbk::Work<double>& work<double>(pb_ctx* ctx)
    {
        if (!ctx->work64)
        {
            allocWork(ctx, ctx->dwork64, ctx->W64);
            if (ctx->rigSys) allocSysWork(ctx, ctx->dwork64, ctx->W64, ctx->Wsys64);
            ctx->work64 = true;
        }
        return ctx->W64;
    }

    // The reduced system the solve routes take: the rig system of a camera rig problem, else the camera-level one.
    const pbk::Problem& sysProblem(const pb_ctx* ctx) { return ctx->rigSys ? ctx->Psys : ctx->P; }
    template <typename T>
    pbk::Work<T>& sysWork(pb_ctx* ctx)
    {
        pbk::Work<T>& W = work<T>(ctx);
        if (!ctx->rigSys) return W;
        if constexpr (std::is_same<T, float>::value)
            return ctx->Wsys32;
        else
            return ctx->Wsys64;
    }

    // Resets the device state (cameras, groups, points) to the uploaded initial state.
    void resetState(pb_ctx* ctx)
    {
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        check(hipMemcpyAsync(P.camR, ctx->in.R.data(), sizeof(double) * ctx->in.R.size(), hipMemcpyHostToDevice, s), "state");
        check(hipMemcpyAsync(P.camC, ctx->in.C.data(), sizeof(double) * ctx->in.C.size(), hipMemcpyHostToDevice, s), "state");
        if (!ctx->prep.gParams.empty())
            check(hipMemcpyAsync(P.groupParams, ctx->prep.gParams.data(), sizeof(double) * ctx->prep.gParams.size(),
                                 hipMemcpyHostToDevice, s),
                  "state");
        if (!ctx->lay.pointX.empty())
            check(hipMemcpyAsync(P.pointX, ctx->lay.pointX.data(), sizeof(double) * ctx->lay.pointX.size(),
                                 hipMemcpyHostToDevice, s),
                  "state");
        if (ctx->rigSys)
        {
            const Input& in = ctx->in;
            check(hipMemcpyAsync(P.capR, in.capR.data(), sizeof(double) * in.capR.size(), hipMemcpyHostToDevice, s), "state");
            check(hipMemcpyAsync(P.capC, in.capC.data(), sizeof(double) * in.capC.size(), hipMemcpyHostToDevice, s), "state");
            check(hipMemcpyAsync(P.senR, in.senR.data(), sizeof(double) * in.senR.size(), hipMemcpyHostToDevice, s), "state");
            check(hipMemcpyAsync(P.senC, in.senC.data(), sizeof(double) * in.senC.size(), hipMemcpyHostToDevice, s), "state");
        }
    }

    // canUseReferenceOnlineSchur (ReferenceSchur.cpp:667-673) for frame cameras without rigs, laser ranges, scale
    // bars, pose priors or a camera plane: the online point Schur (path B) unless laser planes are present or
    // nothing primary is refined; otherwise the assembled damped Schur (path A).
    enum Route
    {
        kRoutePcg,
        kRouteDense,
        kRouteBand
    };
    bool planBand(pb_ctx* ctx);

    bool onlineSchur(const pb_ctx* ctx, const pb_solve_options& o)
    {
        return o.use_online_schur != 0 && ctx->prep.primaryBlocks > 0 && ctx->in.nPlanes == 0 && !ctx->in.rig();
    }

    // Per-solve device settings that follow the options: robust loss, the parameterisation
    // (useReferencePointParameterization, Assembly.cpp:24-29) and the survey-constraint table, rebuilt each solve
    // because its weights and Huber thresholds are options. Constraints are grouped by active track, laser planes
    // before control points and each kind in caller order (Assembly.cpp:656-690); a constraint whose weight
    // product is not positive is left out, as its linearisation would fail (Constraints.cpp:72-76, 94-98).
    void applySolveOptions(pb_ctx* ctx, const pb_solve_options& o)
    {
        pbk::Problem& P = ctx->P;
        ctx->pcgGraphs = o.pcg_graphs != 0;
        P.loss = o.robust_loss;
        P.lossScale = o.robust_loss_scale;
        P.refParam = (o.use_online_schur && ctx->in.nControl == 0 && ctx->in.nPlanes == 0 && !ctx->in.rig()) ? 1 : 0;
        ctx->dcons.release();
        P.nCons = 0;
        P.consOffset = P.consPoint = P.consRows = P.consIndex = nullptr;
        P.consJ = P.consTarget = P.consDelta = nullptr;
        if (ctx->in.cons.empty()) return;
        const Layout& L = ctx->lay;
        struct Key
        {
            int active, plane, index, k;
        };
        std::vector<Key> keys;
        for (int k = 0; k < static_cast<int>(ctx->in.cons.size()); ++k)
        {
            const Input::Constraint& c = ctx->in.cons[k];
            const int a = L.activeIndexOf[c.track];
            const double w = (c.plane ? o.laser_plane_weight : o.control_point_weight) * c.weight;
            if (a < 0 || !(w > 0.0) || !std::isfinite(w)) continue;
            keys.push_back({a, c.plane ? 0 : 1, c.index, k});
        }
        std::sort(keys.begin(), keys.end(), [](const Key& x, const Key& y) {
            return x.active != y.active ? x.active < y.active : x.plane != y.plane ? x.plane < y.plane : x.index < y.index;
        });
        const int n = static_cast<int>(keys.size());
        std::vector<int32_t> offset(L.nActive + 1, 0), point(n), rows(n), index(n);
        std::vector<double> J(size_t(n) * 9, 0.0), target(size_t(n) * 3), delta(n);
        for (int i = 0; i < n; ++i)
        {
            const Input::Constraint& c = ctx->in.cons[keys[i].k];
            const double sw = std::sqrt((c.plane ? o.laser_plane_weight : o.control_point_weight) * c.weight);
            ++offset[keys[i].active + 1];
            point[i] = keys[i].active;
            rows[i] = c.plane ? 1 : 3;
            index[i] = c.plane ? -1 : c.index;
            for (int q = 0; q < (c.plane ? 3 : 9); ++q) J[9 * i + q] = sw * c.A[q];
            std::copy(c.target, c.target + 3, &target[3 * i]);
            delta[i] = c.plane ? o.laser_huber_delta * sw : o.control_point_huber_delta;
        }
        for (int a = 0; a < L.nActive; ++a) offset[a + 1] += offset[a];
        if (n == 0) return;
        hipStream_t s = ctx->stream;
        DeviceAlloc& A = ctx->dcons;
        P.nCons = n;
        P.consOffset = A.upload(offset, s);
        P.consPoint = A.upload(point, s);
        P.consRows = A.upload(rows, s);
        P.consIndex = A.upload(index, s);
        P.consJ = A.upload(J, s);
        P.consTarget = A.upload(target, s);
        P.consDelta = A.upload(delta, s);
    }

    struct BuildResult
    {
        double cost = 0.0;
        bool invalid = false, singular = false, notSpd = false;
    };

    // Enqueues the linearisation, the camera terms and the Schur assembly at damping, and (unless fused, where
    // the step publishes everything at its end and the unit step length is set here) the copies of the
    // objective and the flags to the pinned area (read by buildRead after the stream is synchronised).
    template <typename T>
    void buildLaunch(pb_ctx* ctx, pbk::Work<T>& W, double damping, bool fused = false)
    {
        using Lr = pbk::Launcher<T>;
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        Lr::clearBuild(P, W, fused, s);
        if (ctx->rigSys) Lr::rigMaps(P, s);
        Lr::prepareCameras(P, s);
        {
            Span sp(ctx, 0);
            Lr::linearizePoints(P, W, damping, s);
        }
        {
            Span sp(ctx, 1);
            Lr::cameraDirect(P, W, s);
        }
        {
            Span sp(ctx, 2);
            Lr::schurBlocks(P, W, damping, s);
            if (ctx->rigSys) Lr::rigProject(P, ctx->Psys, W, sysWork<T>(ctx), damping, s);
            if (ctx->cluster.active)
                pbk::ClusterLauncher<T>::factor(sysProblem(ctx), sysWork<T>(ctx).S, ctx->cluster.dev, W.flags, s);
        }
        if (fused) return;
        double* h = ctx->pinned;
        check(hipMemcpyAsync(h, W.scalars, sizeof(double) * pbk::kScalarCount, hipMemcpyDeviceToHost, s), "readback");
        check(hipMemcpyAsync(h + kPinFlags, W.flags, sizeof(int32_t) * pbk::kFlagCount, hipMemcpyDeviceToHost, s),
              "readback");
    }

    BuildResult buildRead(const pb_ctx* ctx)
    {
        const double* h = ctx->pinned;
        const int32_t* f = reinterpret_cast<const int32_t*>(h + kPinFlags);
        BuildResult r;
        r.cost = h[pbk::kScalarCost];
        r.invalid = f[pbk::kFlagInvalid] != 0;
        r.singular = f[pbk::kFlagSingular] != 0;
        r.notSpd = f[pbk::kFlagNotSpd] != 0;
        return r;
    }

    template <typename T>
    BuildResult build(pb_ctx* ctx, pbk::Work<T>& W, double damping)
    {
        buildLaunch(ctx, W, damping);
        check(hipStreamSynchronize(ctx->stream), "build");
        return buildRead(ctx);
    }

    constexpr int kPcgGraphLaunches = 16;  // PCG launches per captured graph (even: the counter stays even)
    static_assert(kPcgGraphLaunches % 2 == 0, "a replay keeps the iteration parity");

    // The captured PCG iterations for these launch arguments (captured on first use and after any change).
    template <typename T>
    hipGraphExec_t pcgGraph(pb_ctx* ctx, pbk::Work<T>& W, int maxIter, double relTol, double absTol, double modelEta)
    {
        using Lr = pbk::Launcher<T>;
        const pbk::Problem& P = sysProblem(ctx);
        hipStream_t s = ctx->stream;
        auto& g = ctx->pcgGraph[std::is_same<T, double>::value ? 1 : 0];
        std::vector<unsigned char> key(sizeof(P) + sizeof(W) + sizeof(int) + 3 * sizeof(double));
        unsigned char* k = key.data();
        std::memcpy(k, &P, sizeof(P));
        k += sizeof(P);
        std::memcpy(k, &W, sizeof(W));
        k += sizeof(W);
        std::memcpy(k, &maxIter, sizeof(int));
        k += sizeof(int);
        const double d[3] = {relTol, absTol, modelEta};
        std::memcpy(k, d, sizeof(d));
        if (g.exec && g.key == key) return g.exec;
        if (g.exec)
        {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        for (int j = 0; j < kPcgGraphLaunches; ++j)
            Lr::pcgIteration(P, W, j, maxIter, relTol, absTol, modelEta, s, true);
        Lr::pcgAdvance(P, W, kPcgGraphLaunches, s);
        check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) g.exec = nullptr;
        check(e, "hipGraphInstantiate");
        g.key = std::move(key);
        return g.exec;
    }

    struct PcgResult
    {
        bool converged = false, breakdown = false, modelStop = false;
        int iterations = 0;
        double rnorm = 0.0;
    };

    // PCG with the cluster block-Jacobi preconditioner (four launches per iteration, chunks between host polls).
    template <typename T>
    PcgResult clusterPcg(pb_ctx* ctx, pbk::Work<T>& W, double relTol, double absTol, int maxIter, int checkEvery,
                         double modelEta)
    {
        using CL = pbk::ClusterLauncher<T>;
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = sysProblem(ctx);
        PcgResult res;
        if (P.nUnits == 0)
        {
            res.converged = true;
            return res;
        }
        double* h = ctx->pinned + kPinPcg;
        int32_t* hf = reinterpret_cast<int32_t*>(ctx->pinned + kPinPcgFlags);
        auto readback = [&]() {
            check(hipMemcpyAsync(hf, W.iflags, sizeof(int32_t) * pbk::kIflagCount, hipMemcpyDeviceToHost, s), "readback");
            check(hipMemcpyAsync(h, W.scalars, sizeof(double) * pbk::kScalarCount, hipMemcpyDeviceToHost, s), "readback");
            check(hipStreamSynchronize(s), "pcg");
        };
        CL::init(P, W, ctx->cluster.dev, relTol, absTol, s);
        int it = 0;
        while (true)
        {
            const int chunk = std::min(std::max(checkEvery, it / 4), 256);
            const int last = std::min(it + chunk, maxIter);
            for (; it < last; ++it) CL::iteration(P, W, ctx->cluster.dev, it, maxIter, modelEta, s);
            readback();
            if (hf[pbk::kIflagDone0] || it >= maxIter) break;
        }
        res.breakdown = hf[pbk::kIflagBreakdown] != 0;
        res.iterations = hf[pbk::kIflagIters];
        res.rnorm = h[pbk::kScalarRnorm];
        res.modelStop = hf[pbk::kIflagModelStop] != 0;
        const bool truncated = modelEta > 0.0 && res.iterations >= maxIter;
        res.converged = !res.breakdown && (res.modelStop || truncated || res.rnorm <= h[pbk::kScalarThreshold]);
        if (ctx->profile) ctx->pcgIterationsProfiled += res.iterations;
        return res;
    }

    template <typename T>
    PcgResult pcg(pb_ctx* ctx, pbk::Work<T>& W, double relTol, double absTol, int maxIter, int checkEvery,
                  double modelEta)
    {
        if (ctx->cluster.active) return clusterPcg(ctx, W, relTol, absTol, maxIter, checkEvery, modelEta);
        using Lr = pbk::Launcher<T>;
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = sysProblem(ctx);
        PcgResult res;
        double* h = ctx->pinned + kPinPcg;
        int32_t* hf = reinterpret_cast<int32_t*>(ctx->pinned + kPinPcgFlags);
        auto readback = [&]() {
            check(hipMemcpyAsync(hf, W.iflags, sizeof(int32_t) * pbk::kIflagCount, hipMemcpyDeviceToHost, s), "readback");
            check(hipMemcpyAsync(h, W.scalars, sizeof(double) * pbk::kScalarCount, hipMemcpyDeviceToHost, s), "readback");
            check(hipStreamSynchronize(s), "pcg");
        };
        if (P.cgBlocks == 0)  // no primary unknowns: the zero step solves the empty system
        {
            res.converged = true;
            return res;
        }
        // fused iterations; the host looks at the flags after chunks that grow with the iteration count
        // (launches after the stop return at once)
        // with pcg_graphs, whole replays of the captured launches (a launch past the stop, or at maxIter where
        // the kernel stops, returns at once)
        hipGraphExec_t exec = ctx->pcgGraphs ? pcgGraph<T>(ctx, W, maxIter, relTol, absTol, modelEta) : nullptr;
        Lr::pcgInit(P, W, relTol, absTol, s);
        int it = 0;
        while (true)
        {
            const int chunk = std::min(std::max(checkEvery, it / 4), 256);
            const int last = std::min(it + chunk, maxIter + 1);
            const int first = it;
            TimedSpan ts{nullptr, nullptr, 0};
            if (ctx->profile)
            {
                ts.start = takeEvent(ctx);
                ts.stop = takeEvent(ctx);
                check(hipEventRecord(ts.start, s), "hipEventRecord");
            }
            if (exec)
                for (; it < last; it += kPcgGraphLaunches) check(hipGraphLaunch(exec, s), "hipGraphLaunch");
            else
                for (; it < last; ++it) Lr::pcgIteration(P, W, it, maxIter, relTol, absTol, modelEta, s);
            ts.launches = it - first;
            if (ctx->profile) check(hipEventRecord(ts.stop, s), "hipEventRecord");
            readback();
            const bool done = hf[pbk::kIflagDone0] || hf[pbk::kIflagDone1] || it > maxIter;
            if (ctx->profile && !done) ctx->fam[3].pending.push_back(ts);
            if (done) break;
        }
        res.breakdown = hf[pbk::kIflagBreakdown] != 0;
        res.iterations = hf[pbk::kIflagIters];
        res.rnorm = h[pbk::kScalarRnorm];
        res.modelStop = hf[pbk::kIflagModelStop] != 0;
        // with the truncated-Newton test on, an iterate cut at maxIter is a step too: every CG iterate from zero
        // decreases the quadratic model, so it is a descent direction for the line search
        const bool truncated = modelEta > 0.0 && res.iterations >= maxIter;
        res.converged = !res.breakdown && (res.modelStop || truncated || res.rnorm <= h[pbk::kScalarThreshold]);
        if (ctx->profile) ctx->pcgIterationsProfiled += res.iterations;
        return res;
    }

    // ------------------------------------------------------------------------------------------
    // Dense reduced solve (pb_dense.hip): Cholesky of the gathered reduced system in FP64, or a guarded FP32
    // factorisation whose solution is refined in FP64 against the stored system.
    int denseTiles(const pbk::Problem& P) { return (6 * P.nUnits + pbk::kDenseTile - 1) / pbk::kDenseTile; }

    // Vectors and state words of a direct reduced solve; with matrix, the dense matrix of precision fp32/fp64 too
    // (the band route keeps its factors in its own arena).
    void ensureDense(pb_ctx* ctx, int nt, bool fp32, bool matrix = true)
    {
        auto& D = ctx->dense;
        hipStream_t s = ctx->stream;
        const size_t npad = size_t(nt) * pbk::kDenseTile;
        if (nt > D.ntCap)
        {
            D.vec.release();
            D.mat32.release();
            D.mat64.release();
            D.A32 = nullptr;
            D.A64 = nullptr;
            D.ntCap = nt;
            D.v32 = D.vec.alloc<float>(npad);
            D.v64 = D.vec.alloc<double>(npad);
            D.x = D.vec.alloc<double>(npad);
            D.r = D.vec.alloc<double>(npad);
            D.rowNorm = D.vec.alloc<double>(npad);
            D.norm = D.vec.alloc<double>(4);
            D.st.words = D.vec.alloc<int32_t>(pbk::kDenseWordCount);
            D.st.progress = D.vec.alloc<unsigned long long>(2);
            // words and progress start at zero once; every later solve advances the epoch on the device
            check(hipMemsetAsync(D.st.words, 0, sizeof(int32_t) * pbk::kDenseWordCount, s), "memset");
            check(hipMemsetAsync(D.st.progress, 0, sizeof(unsigned long long) * 2, s), "memset");
        }
        if (!matrix) return;
        const size_t cap = size_t(D.ntCap) * pbk::kDenseTile;
        if (fp32 && !D.A32) D.A32 = D.mat32.alloc<float>(cap * cap);
        if (!fp32 && !D.A64) D.A64 = D.mat64.alloc<double>(cap * cap);
    }

    // ------------------------------------------------------------------------------------------
    // Band route (pb_band.hip): plan of the block-tridiagonal cyclic reduction for the current pattern.

    // Reverse Cuthill-McKee order of an undirected graph (no self loops): per connected component a
    // breadth-first search from a pseudo-peripheral vertex (repeated searches from the least-degree vertex of
    // the last level while the eccentricity grows), neighbours visited by increasing degree, then the whole
    // order reversed.
    std::vector<int> reverseCuthillMcKee(const std::vector<std::vector<int>>& adj)
    {
        const int n = static_cast<int>(adj.size());
        std::vector<int> order;
        order.reserve(n);
        std::vector<char> placed(n, 0);
        std::vector<int> mark(n, -1);
        int stamp = 0;
        // vertices of the component of root by levels; returns the last level
        auto levels = [&](int root, int& depth) {
            ++stamp;
            std::vector<int> cur{root}, last;
            mark[root] = stamp;
            depth = 0;
            while (!cur.empty())
            {
                last = cur;
                std::vector<int> next;
                for (int v : cur)
                    for (int w : adj[v])
                        if (mark[w] != stamp && !placed[w])
                        {
                            mark[w] = stamp;
                            next.push_back(w);
                        }
                if (!next.empty()) ++depth;
                cur.swap(next);
            }
            return last;
        };
        auto leastDegree = [&](const std::vector<int>& vs) {
            int best = vs.front();
            for (int v : vs)
                if (adj[v].size() < adj[best].size() || (adj[v].size() == adj[best].size() && v < best)) best = v;
            return best;
        };
        for (int s0 = 0; s0 < n; ++s0)
        {
            if (placed[s0]) continue;
            int root = s0, depth = 0;
            std::vector<int> last = levels(root, depth);
            for (int round = 0; round < 8; ++round)
            {
                const int cand = leastDegree(last);
                int d2 = 0;
                std::vector<int> l2 = levels(cand, d2);
                if (d2 <= depth) break;
                root = cand;
                depth = d2;
                last.swap(l2);
            }
            // Cuthill-McKee from root
            size_t head = order.size();
            order.push_back(root);
            placed[root] = 1;
            while (head < order.size())
            {
                const int v = order[head++];
                std::vector<int> nb;
                for (int w : adj[v])
                    if (!placed[w]) nb.push_back(w);
                std::sort(nb.begin(), nb.end(), [&](int x, int y) {
                    return adj[x].size() != adj[y].size() ? adj[x].size() < adj[y].size() : x < y;
                });
                nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
                for (int w : nb)
                    if (!placed[w])
                    {
                        placed[w] = 1;
                        order.push_back(w);
                    }
            }
        }
        std::reverse(order.begin(), order.end());
        return order;
    }

    // Plans the band route for the current pattern (once per pattern build). The pose units are ordered by
    // reverse Cuthill-McKee over the pose-pose blocks of S, the bandwidth w (in units) fixes the super-block
    // order sp = 64 ceil(6 w / 64) <= 64 kBandMaxTiles, and each super-block takes floor(sp / 6) units, so no
    // stored block couples super-blocks further apart than one. The group units form the border. The cyclic
    // reduction schedule (levels, coupling slots, tasks) follows; every arena offset is checked against the
    // arena size here, so the kernels index only inside it.
    bool planBand(pb_ctx* ctx)
    {
        auto& Bd = ctx->band;
        if (Bd.planBuild == ctx->patternBuilds) return Bd.ok;
        Bd.tables.release();
        Bd.arena32.release();
        Bd.arena64.release();
        Bd.Z32 = nullptr;
        Bd.Z64 = nullptr;
        Bd.view = pbk::BandView{};
        Bd.ok = false;
        Bd.bandwidth = 0;
        Bd.planBuild = ctx->patternBuilds;
        const Layout& L = ctx->lay;
        const int U = ctx->prep.nUnits, nPose = ctx->prep.cameraBlocks;
        if (nPose < 2)
        {
            Bd.why = "fewer than two pose units";
            return false;
        }
        std::vector<std::vector<int>> adj(nPose);
        for (int u = 0; u < nPose; ++u)
            for (int k = L.rows.rowPtr[u]; k < L.rows.rowPtr[u + 1]; ++k)
            {
                const int v = L.rows.colIdx[k];
                if (v < nPose && v != u) adj[u].push_back(v);
            }
        const std::vector<int> order = reverseCuthillMcKee(adj);
        std::vector<int> pos(nPose);
        for (int i = 0; i < nPose; ++i) pos[order[i]] = i;
        int w = 1;
        for (int u = 0; u < nPose; ++u)
            for (int v : adj[u]) w = std::max(w, std::abs(pos[u] - pos[v]));
        Bd.bandwidth = w;
        const int t = (6 * w + pbk::kDenseTile - 1) / pbk::kDenseTile;
        if (t > pbk::kBandMaxTiles)
        {
            Bd.why = "bandwidth of " + std::to_string(w) + " pose units exceeds " +
                     std::to_string(pbk::kBandMaxRows / 6);
            return false;
        }
        pbk::BandView V;
        pbk::BandDev& D = V.d;
        D.sp = pbk::kDenseTile * t;
        D.su = D.sp / 6;
        D.nPose = nPose;
        D.m = (nPose + D.su - 1) / D.su;
        D.nb = 6 * (U - nPose);
        D.nbp = (D.nb + pbk::kDenseTile - 1) / pbk::kDenseTile * pbk::kDenseTile;
        if (D.nbp > pbk::kBandMaxBorder)
        {
            Bd.why = "border of " + std::to_string(D.nb) + " intrinsic rows exceeds " +
                     std::to_string(pbk::kBandMaxBorder);
            return false;
        }
        if (D.m < 2)
        {
            Bd.why = "a single super-block";
            return false;
        }
        const int sp = D.sp, m = D.m, nbp = D.nbp, tb = nbp / pbk::kDenseTile;
        D.nF = sp + nbp;
        D.n = 6 * U;
        D.vN = m * sp + nbp;
        // unit placement
        std::vector<int32_t> unitBlk(U), unitOff(U), rowOf(D.n), srcOf(D.vN, -1);
        for (int u = 0; u < U; ++u)
        {
            if (u < nPose)
            {
                unitBlk[u] = pos[u] / D.su;
                unitOff[u] = 6 * (pos[u] % D.su);
            }
            else
            {
                unitBlk[u] = -1;
                unitOff[u] = 6 * (u - nPose);
            }
            const int base = u < nPose ? unitBlk[u] * sp + unitOff[u] : m * sp + unitOff[u];
            for (int a = 0; a < 6; ++a)
            {
                rowOf[6 * u + a] = base + a;
                srcOf[base + a] = 6 * u + a;
            }
        }
        // the pattern must be block-tridiagonal over the super-blocks
        for (int u = 0; u < nPose; ++u)
            for (int v : adj[u])
                if (std::abs(unitBlk[u] - unitBlk[v]) > 1) throw std::logic_error("band plan: coupling beyond a neighbour");
        // cyclic reduction schedule; a slot couples two consecutive active super-blocks, stored with the rows of
        // the one eliminated at that level
        struct Slot
        {
            int rows, cols;
        };
        std::vector<Slot> slots;
        std::vector<int32_t> k0slot(m - 1);
        std::vector<int> act(m), actSlot(m - 1);  // active blocks and the slots between neighbours
        for (int k = 0; k < m; ++k) act[k] = k;
        for (int p = 0; p + 1 < m; ++p)
        {
            const bool oddFirst = (p & 1) != 0;  // positions p (odd: eliminated) and p + 1
            slots.push_back(oddFirst ? Slot{p, p + 1} : Slot{p + 1, p});
            k0slot[p] = p;
            actSlot[p] = p;
        }
        struct ElimRec
        {
            int e, a, b, sA, sB;
        };
        std::vector<std::vector<ElimRec>> levelElims;
        std::vector<pbk::BandTask> tasks;
        std::vector<pbk::BandTile> tiles;
        V.tileXStart.push_back(0);
        V.tileUStart.push_back(0);
        const long long spp = (long long)sp * sp, spb = (long long)sp * nbp, bb = (long long)nbp * nbp;
        // arena offsets are fixed once the slot count is known; tasks are written with symbolic slot ids first
        // and resolved below
        struct TaskSym
        {
            int kind;  // 0: X of slot, 1: XB of block, 2: D of survivor, 3: BdT of survivor, 4: new slot, 5: Pc
            int a0, a1, a2, a3, a4, a5;
            int terms;
        };
        std::vector<std::vector<TaskSym>> symX, symU;
        while (act.size() > 1)
        {
            const int na = static_cast<int>(act.size());
            std::vector<ElimRec> el;
            std::vector<TaskSym> sx, su;
            for (int j = 0; 2 * j + 1 < na; ++j)
            {
                ElimRec r;
                r.e = act[2 * j + 1];
                r.a = act[2 * j];
                r.sA = actSlot[2 * j];
                if (2 * j + 2 < na)
                {
                    r.b = act[2 * j + 2];
                    r.sB = actSlot[2 * j + 1];
                }
                else
                {
                    r.b = -1;
                    r.sB = -1;
                }
                el.push_back(r);
                sx.push_back({0, r.sA, r.e, 0, 0, 0, 0, 1});
                if (r.sB >= 0) sx.push_back({0, r.sB, r.e, 0, 0, 0, 0, 1});
                if (nbp > 0) sx.push_back({1, r.e, 0, 0, 0, 0, 0, 1});
            }
            std::vector<int> next, nextSlot;
            for (int q = 0; 2 * q < na; ++q) next.push_back(act[2 * q]);
            for (int q = 0; q < static_cast<int>(next.size()); ++q)
            {
                const int s = next[q];
                const bool hasL = q > 0, hasR = 2 * q + 1 < na;
                const int eL = hasL ? act[2 * q - 1] : -1, sL = hasL ? actSlot[2 * q - 1] : -1;
                const int eR = hasR ? act[2 * q + 1] : -1, sR = hasR ? actSlot[2 * q] : -1;
                if (!hasL && !hasR) continue;
                const int terms = (hasL ? 1 : 0) + (hasR ? 1 : 0);
                const int s0 = hasL ? sL : sR, e0 = hasL ? eL : eR;
                const int s1 = hasL && hasR ? sR : -1, e1 = hasL && hasR ? eR : -1;
                su.push_back({2, s, s0, s1, 0, 0, 0, terms});
                if (nbp > 0) su.push_back({3, s, s0, e0, s1, e1, 0, terms});
            }
            for (int q = 0; q + 1 < static_cast<int>(next.size()); ++q)
            {
                const int xa = actSlot[2 * q], xb = actSlot[2 * q + 1];
                const int id = static_cast<int>(slots.size());
                if (q % 2 == 0)
                {
                    slots.push_back({next[q + 1], next[q]});  // eliminated next: next[q + 1]
                    su.push_back({4, id, xb, xa, 0, 0, 0, 1});
                }
                else
                {
                    slots.push_back({next[q], next[q + 1]});
                    su.push_back({4, id, xa, xb, 0, 0, 0, 1});
                }
                nextSlot.push_back(id);
            }
            if (nbp > 0)
                for (const ElimRec& r : el) su.push_back({5, r.e, 0, 0, 0, 0, 0, 1});
            levelElims.push_back(el);
            symX.push_back(sx);
            symU.push_back(su);
            act.swap(next);
            actSlot.swap(nextSlot);
        }
        V.levels = static_cast<int>(levelElims.size());
        D.nSlots = static_cast<int>(slots.size());
        auto align = [](long long x) { return (x + 63) / 64 * 64; };
        D.offD = 0;
        D.offK = align(D.offD + m * spp);
        D.offB = align(D.offK + D.nSlots * spp);
        D.offC = align(D.offB + m * spb);
        D.offX = align(D.offC + bb);
        D.offXB = align(D.offX + D.nSlots * spp);
        D.offPc = align(D.offXB + m * spb);
        D.offF = align(D.offPc + m * bb);
        D.offV = align(D.offF + (long long)D.nF * D.nF);
        D.offFV = align(D.offV + D.vN);
        D.offPV = align(D.offFV + D.nF);
        D.offPB = align(D.offPV + (long long)D.nSlots * sp);
        D.offY = align(D.offPB + (long long)m * nbp);
        D.size = align(D.offY + (long long)m * sp);
        // tasks and tiles
        auto addTask = [&](pbk::BandTask tk, int rowTiles, int colTiles, bool lower, long long outExtent,
                           std::vector<pbk::BandTile>& out) {
            // every element a tile touches lies inside the arena
            auto inside = [&](long long off, long long extent) {
                if (off < 0 || off + extent > D.size) throw std::logic_error("band plan: task outside the arena");
            };
            inside(tk.out, outExtent);
            for (int q = 0; q < tk.terms; ++q)
            {
                inside(q ? tk.p1 : tk.p0, (long long)rowTiles * pbk::kDenseTile * sp);
                inside(q ? tk.q1 : tk.q0, (long long)colTiles * pbk::kDenseTile * sp);
            }
            const int id = static_cast<int>(tasks.size());
            tasks.push_back(tk);
            for (int tj = 0; tj < colTiles; ++tj)
                for (int ti = 0; ti < rowTiles; ++ti)
                    if (!lower || ti >= tj) out.push_back({id, ti, tj, 0});
        };
        auto X = [&](int slot) { return D.offX + slot * spp; };
        auto XB = [&](int blk) { return D.offXB + blk * spb; };
        for (int l = 0; l < V.levels; ++l)
        {
            for (const TaskSym& ts : symX[l])
            {
                // kind 0: a0 = slot, a1 = eliminated block; kind 1: a0 = eliminated block
                pbk::BandTask tk{};
                tk.terms = 1;
                tk.kTri = 1;
                tk.ldo = sp;
                if (ts.kind == 0)
                {
                    tk.p0 = D.offD + ts.a1 * spp;  // W_e
                    tk.out = X(ts.a0);
                    tk.q0 = D.offK + ts.a0 * spp;
                    addTask(tk, t, t, false, spp, tiles);
                }
                else
                {
                    tk.p0 = D.offD + ts.a0 * spp;
                    tk.out = XB(ts.a0);
                    tk.q0 = D.offB + ts.a0 * spb;
                    addTask(tk, t, tb, false, spb, tiles);
                }
            }
            V.tileXStart.push_back(static_cast<int>(tiles.size()));
            V.tileUStart.push_back(0);
        }
        // the update tiles follow all X tiles in the tile array
        V.tileUStart[0] = static_cast<int>(tiles.size());
        for (int l = 0; l < V.levels; ++l)
        {
            for (const TaskSym& ts : symU[l])
            {
                pbk::BandTask tk{};
                tk.terms = ts.terms;
                tk.ldo = sp;
                switch (ts.kind)
                {
                case 2:  // D_s -= sum X^T X
                    tk.out = D.offD + ts.a0 * spp;
                    tk.p0 = X(ts.a1);
                    tk.q0 = X(ts.a1);
                    if (ts.terms > 1)
                    {
                        tk.p1 = X(ts.a2);
                        tk.q1 = X(ts.a2);
                    }
                    tk.alphaNeg = 1;
                    tk.beta = 1;
                    addTask(tk, t, t, true, spp, tiles);
                    break;
                case 3:  // BdT_s -= sum X^T XB
                    tk.out = D.offB + ts.a0 * spb;
                    tk.p0 = X(ts.a1);
                    tk.q0 = XB(ts.a2);
                    if (ts.terms > 1)
                    {
                        tk.p1 = X(ts.a3);
                        tk.q1 = XB(ts.a4);
                    }
                    tk.alphaNeg = 1;
                    tk.beta = 1;
                    addTask(tk, t, tb, false, spb, tiles);
                    break;
                case 4:  // new coupling = -P^T Q
                    tk.out = D.offK + ts.a0 * spp;
                    tk.p0 = X(ts.a1);
                    tk.q0 = X(ts.a2);
                    tk.alphaNeg = 1;
                    addTask(tk, t, t, false, spp, tiles);
                    break;
                default:  // Pc_e = XB^T XB
                    tk.out = D.offPc + ts.a0 * bb;
                    tk.ldo = nbp;
                    tk.p0 = XB(ts.a0);
                    tk.q0 = XB(ts.a0);
                    addTask(tk, tb, tb, false, bb, tiles);
                    break;
                }
            }
            V.tileUStart[l + 1] = static_cast<int>(tiles.size());
        }
        // pending slots per block (forward sweep): the slots whose columns are the block, in creation order
        std::vector<std::vector<int>> pendOf(m);
        for (int sId = 0; sId < D.nSlots; ++sId) pendOf[slots[sId].cols].push_back(sId);
        std::vector<int32_t> pend;
        std::vector<pbk::BandElim> elims;
        for (int l = 0; l < V.levels; ++l)
        {
            V.elimStart.push_back(static_cast<int>(elims.size()));
            for (const ElimRec& r : levelElims[l])
            {
                pbk::BandElim be{};
                be.e = r.e;
                be.a = r.a;
                be.b = r.b;
                be.slotA = r.sA;
                be.slotB = r.sB;
                be.pendBegin = static_cast<int32_t>(pend.size());
                for (int sId : pendOf[r.e]) pend.push_back(sId);
                be.pendEnd = static_cast<int32_t>(pend.size());
                if (slots[r.sA].rows != r.e || (r.sB >= 0 && slots[r.sB].rows != r.e))
                    throw std::logic_error("band plan: slot orientation");
                elims.push_back(be);
            }
        }
        V.elimStart.push_back(static_cast<int>(elims.size()));
        D.lastPendBegin = static_cast<int32_t>(pend.size());
        for (int sId : pendOf[0]) pend.push_back(sId);
        D.lastPendEnd = static_cast<int32_t>(pend.size());
        if (static_cast<int>(elims.size()) != m - 1) throw std::logic_error("band plan: every block but one eliminated");
        // every index the kernels follow stays inside its table or the arena
        for (const pbk::BandElim& be : elims)
            if (be.e <= 0 || be.e >= m || be.a < 0 || be.a >= m || be.b < -1 || be.b >= m || be.slotA < 0 ||
                be.slotA >= D.nSlots || be.slotB < -1 || be.slotB >= D.nSlots || (be.b < 0) != (be.slotB < 0))
                throw std::logic_error("band plan: elimination record out of range");
        for (int32_t sId : pend)
            if (sId < 0 || sId >= D.nSlots) throw std::logic_error("band plan: pending slot out of range");
        for (int u = 0; u < U; ++u)
            if (unitBlk[u] < -1 || unitBlk[u] >= m || unitOff[u] < 0 ||
                unitOff[u] + 6 > (unitBlk[u] >= 0 ? sp : D.nb))
                throw std::logic_error("band plan: unit placement out of range");
        // algorithmic flops: per eliminated block a Cholesky and a triangular inverse of order sp (sp^3 / 3
        // each), every task tile 2 * 64 * 64 * k per term, the border sums, the last dense system (nF^3 / 3);
        // the solve: W^T r and W t, the coupling and border products of both sweeps, the last system's sweeps
        {
            const double spd = sp, nbd = nbp, nFd = D.nF;
            double f = (m - 1) * (2.0 * spd * spd * spd / 3.0) + (m - 1) * nbd * nbd + nFd * nFd * nFd / 3.0;
            for (const pbk::BandTile& tl : tiles)
            {
                const pbk::BandTask& tk = tasks[tl.task];
                const double kEnd = tk.kTri ? std::min(spd, (tl.ti + 1) * 64.0) : spd;
                f += 2.0 * 64.0 * 64.0 * kEnd * tk.terms;
            }
            V.factorFlops = f;
            double g = 2.0 * nFd * nFd;
            for (const pbk::BandElim& be : elims)
                g += 2.0 * (2.0 * spd * spd / 2.0 + (be.slotB >= 0 ? 4.0 : 2.0) * spd * spd + 2.0 * spd * nbd);
            V.solveFlops = g;
        }
        hipStream_t s = ctx->stream;
        D.unitBlk = Bd.tables.upload(unitBlk, s);
        D.unitOff = Bd.tables.upload(unitOff, s);
        D.k0slot = Bd.tables.upload(k0slot, s);
        D.rowOf = Bd.tables.upload(rowOf, s);
        D.srcOf = Bd.tables.upload(srcOf, s);
        D.elims = Bd.tables.upload(elims, s);
        D.pend = Bd.tables.upload(pend, s);
        D.tasks = Bd.tables.upload(tasks, s);
        D.tiles = Bd.tables.upload(tiles, s);
        check(hipStreamSynchronize(s), "band plan upload");
        Bd.view = std::move(V);
        Bd.ok = true;
        Bd.why.clear();
        return true;
    }

    // Reduced-solve route of a solve. PB_REDUCED_AUTO takes the band route when the pattern is block-tridiagonal
    // over at least three super-blocks, else the dense Cholesky up to dense_max_rows rows, else PCG.
    // Clusters of `size` consecutive primary blocks of the reduced system the routes solve (a pose unit, or all
    // units of one intrinsic group), as contiguous unit ranges. The cluster PCG runs when some cluster spans more
    // than one unit; otherwise the fused PCG's per-unit blocks are that preconditioner.
    bool planClusters(pb_ctx* ctx, int size)
    {
        auto& Cl = ctx->cluster;
        if (Cl.planBuild == ctx->patternBuilds && Cl.size == size)
        {
            Cl.active = Cl.wide;
            return Cl.active;
        }
        Cl.mem.release();
        Cl.dev = pbk::ClusterDev{};
        Cl.active = Cl.wide = false;
        Cl.size = size;
        Cl.planBuild = ctx->patternBuilds;
        const Prepared& p = ctx->prep;
        const int pose = p.refCameraBlocks();
        std::vector<std::pair<int, int>> primary;  // unit ranges of the primary blocks, in order
        for (int u = 0; u < pose; ++u) primary.push_back({u, u + 1});
        for (int g = 0; g < p.G; ++g)
            if (p.gK[g] > 0)
            {
                const int first = p.gUnit[g] - p.cameraBlocks + pose;
                primary.push_back({first, first + (p.gK[g] + 5) / 6});
            }
        std::vector<int32_t> begin{0}, unitCluster;
        std::vector<long long> off{0};
        bool wide = false;
        int maxDim = 0;
        for (size_t b = 0; b < primary.size(); b += size)
        {
            const int u0 = primary[b].first, u1 = primary[std::min(primary.size(), b + size) - 1].second;
            wide = wide || u1 - u0 > 1;
            const int d = 6 * (u1 - u0);
            maxDim = std::max(maxDim, d);
            for (int u = u0; u < u1; ++u) unitCluster.push_back(static_cast<int32_t>(begin.size() - 1));
            begin.push_back(u1);
            off.push_back(off.back() + static_cast<long long>(d) * d);
        }
        if (!wide) return false;
        hipStream_t s = ctx->stream;
        pbk::ClusterDev& D = Cl.dev;
        D.n = static_cast<int>(begin.size()) - 1;
        D.maxDim = maxDim;
        D.unitBegin = Cl.mem.upload(begin, s);
        D.unitCluster = Cl.mem.upload(unitCluster, s);
        D.matOff = Cl.mem.upload(off, s);
        D.inv = Cl.mem.alloc<double>(static_cast<size_t>(off.back()));
        D.work = Cl.mem.alloc<double>(static_cast<size_t>(off.back()));
        check(hipStreamSynchronize(s), "cluster plan");
        Cl.active = Cl.wide = true;
        return true;
    }

    Route reducedRoute(pb_ctx* ctx, const pb_solve_options& o)
    {
        if (o.reduced_solver == PB_REDUCED_PCG) return kRoutePcg;
        if (o.reduced_solver == PB_REDUCED_DENSE) return kRouteDense;
        if (ctx->rigSys)
        {
            // the band plan orders camera-level pose units; a rig system takes the dense or the PCG route
            if (o.reduced_solver == PB_REDUCED_BAND)
                throw std::invalid_argument("unsupported: the band route does not take a camera rig's system");
            const int maxRows = o.dense_max_rows > 0 ? o.dense_max_rows : kDenseSolveMaxRows;
            return 6L * ctx->Psys.nUnits <= maxRows ? kRouteDense : kRoutePcg;
        }
        const bool band = planBand(ctx);
        if (o.reduced_solver == PB_REDUCED_BAND)
        {
            if (!band) throw std::invalid_argument("unsupported: the band route does not fit this reduced system: " + ctx->band.why);
            return kRouteBand;
        }
        if (band && ctx->band.view.d.m >= 3) return kRouteBand;
        const int maxRows = o.dense_max_rows > 0 ? o.dense_max_rows : kDenseSolveMaxRows;
        return 6L * ctx->P.nUnits <= maxRows ? kRouteDense : kRoutePcg;
    }

    // Arena of the band route in the factor precision (allocated once per plan and precision).
    template <typename TA>
    TA* bandArena(pb_ctx* ctx)
    {
        auto& Bd = ctx->band;
        const size_t n = static_cast<size_t>(Bd.view.d.size);
        if constexpr (std::is_same<TA, float>::value)
        {
            if (!Bd.Z32) Bd.Z32 = Bd.arena32.alloc<float>(n);
            return Bd.Z32;
        }
        else
        {
            if (!Bd.Z64) Bd.Z64 = Bd.arena64.alloc<double>(n);
            return Bd.Z64;
        }
    }

    struct DenseOutcome
    {
        bool factored = false;   // the Cholesky factorisation found every pivot positive
        bool converged = false;  // the refined relative residual reached the tolerance
        int refinements = 0;
        double relResidual = 0.0;
    };

    // Factors S (the work's reduced matrix, gathered in precision TA) and solves S x = rhs into dense.x (FP64):
    // one solve, then up to maxRefine refinement steps x += (L L^T)^-1 (rhs - S x) while the relative residual
    // is above relTol.
    template <typename TA, typename T>
    DenseOutcome denseAttempt(pb_ctx* ctx, const pbk::Work<T>& W, double relTol, int maxRefine, pb_solve_report* rep)
    {
        using DL = pbk::DenseLauncher<TA>;
        using BL = pbk::BandLauncher<TA>;
        const pbk::Problem& P = sysProblem(ctx);
        hipStream_t s = ctx->stream;
        const bool band = ctx->band.active;
        const pbk::BandView& BV = ctx->band.view;
        const int nt = denseTiles(P);
        const int n = 6 * P.nUnits, npad = nt * pbk::kDenseTile;
        ensureDense(ctx, nt, std::is_same<TA, float>::value, !band);
        auto& D = ctx->dense;
        TA* A = nullptr;
        TA* v = nullptr;
        TA* Z = band ? bandArena<TA>(ctx) : nullptr;
        if constexpr (std::is_same<TA, float>::value)
        {
            A = D.A32;
            v = D.v32;
        }
        else
        {
            A = D.A64;
            v = D.v64;
        }
        // one solve of (L L^T) c = src into D.x (overwrite) or added to it
        auto solveInto = [&](const double* src, bool overwrite) {
            Span sp(ctx, band ? kFamilyBandSolve : kFamilyDenseSolve);
            if (band)
            {
                BL::toBand(BV, src, Z, s);
                BL::solve(BV, Z, D.st, s);
                BL::fromBand(BV, D.x, Z, overwrite, s);
            }
            else
            {
                DL::toDense(src, n, npad, v, s);
                DL::solve(A, nt, v, D.st, s);
                DL::accumulate(D.x, v, n, overwrite, s);
            }
        };
        DenseOutcome out;
        auto t0 = Clock::now();
        {
            Span sp(ctx, band ? kFamilyBandFactor : kFamilyDenseFactor);
            if (band)
            {
                BL::template gather<T>(P, W.S, BV, Z, s);
                BL::factor(BV, Z, D.st, s);
            }
            else
            {
                DL::template gather<T>(P, W.S, A, nt, s);
                DL::factor(A, nt, D.st, s);
            }
        }
        DL::norm2(W.rhs, n, D.norm + 1, s);
        int32_t* hw = reinterpret_cast<int32_t*>(ctx->pinned + kPinDense);
        double* hn = ctx->pinned + kPinDense + 4;
        auto readWords = [&]() {
            check(hipMemcpyAsync(hw, D.st.words, sizeof(int32_t) * pbk::kDenseWordCount, hipMemcpyDeviceToHost, s),
                  "readback");
            check(hipMemcpyAsync(hn, D.norm, sizeof(double) * 2, hipMemcpyDeviceToHost, s), "readback");
            check(hipStreamSynchronize(s), "dense solve");
            if (hw[pbk::kDenseTimeout]) throw std::runtime_error("dense triangular solve stopped waiting for a tile");
        };
        readWords();
        if (rep) rep->cholesky_factorization_seconds += seconds(t0);
        if (hw[pbk::kDenseNotSpd]) return out;
        out.factored = true;
        const double bn = std::sqrt(hn[1]);
        auto ts = Clock::now();
        solveInto(W.rhs, true);
        if (rep)
        {
            check(hipStreamSynchronize(s), "dense solve");
            rep->triangular_solve_seconds += seconds(ts);
        }
        double previous = kInf;
        for (int it = 0;; ++it)
        {
            auto tr = Clock::now();
            DL::template residual<T>(P, W.S, W.rhs, D.x, D.r, D.rowNorm, D.norm, s);
            readWords();
            if (rep) rep->residual_check_seconds += seconds(tr);
            out.relResidual = bn > 0.0 ? std::sqrt(hn[0]) / bn : 0.0;
            if (!std::isfinite(out.relResidual)) return out;
            // refinement that stops contracting (an FP32 factor of a system whose conditioning exceeds what FP32
            // resolves) ends the attempt
            const bool stalled = it > 0 && out.relResidual > 0.5 * previous;
            previous = out.relResidual;
            if (out.relResidual <= relTol || it >= maxRefine || stalled)
            {
                out.converged = out.relResidual <= relTol;
                break;
            }
            auto ts2 = Clock::now();
            solveInto(D.r, false);
            if (rep)
            {
                check(hipStreamSynchronize(s), "dense solve");
                rep->triangular_solve_seconds += seconds(ts2);
            }
            ++out.refinements;
        }
        return out;
    }

    // The FP64 attempt without refinement (the route of an FP64 step), enqueued without a host wait: the
    // factorisation, one solve of the reduced right-hand side, the residual and the copy of the solution to W.x;
    // the caller publishes the solver words and the two norms to the pinned area for denseRead. The solve runs whatever
    // the factorisation found (a pivot that is not positive is replaced by 1, so it completes); a step whose
    // factorisation failed is discarded by the caller. ev[0], ev[1] and ev[2] mark the ends of the factorisation,
    // the solve and the residual check.
    template <typename T>
    void denseEnqueue(pb_ctx* ctx, pbk::Work<T>& W, const hipEvent_t* ev)
    {
        using DL = pbk::DenseLauncher<double>;
        using BL = pbk::BandLauncher<double>;
        const pbk::Problem& P = sysProblem(ctx);
        hipStream_t s = ctx->stream;
        const bool band = ctx->band.active;
        const pbk::BandView& BV = ctx->band.view;
        const int nt = denseTiles(P);
        const int n = 6 * P.nUnits, npad = nt * pbk::kDenseTile;
        ensureDense(ctx, nt, false, !band);
        auto& D = ctx->dense;
        double* A = D.A64;
        double* v = D.v64;
        double* Z = band ? bandArena<double>(ctx) : nullptr;
        {
            Span sp(ctx, band ? kFamilyBandFactor : kFamilyDenseFactor);
            if (band)
            {
                BL::template gather<T>(P, W.S, BV, Z, s);
                BL::factor(BV, Z, D.st, s);
            }
            else
            {
                DL::template gather<T>(P, W.S, A, nt, s);
                DL::factor(A, nt, D.st, s);
            }
        }
        DL::norm2(W.rhs, n, D.norm + 1, s);
        check(hipEventRecord(ev[0], s), "event");
        {
            Span sp(ctx, band ? kFamilyBandSolve : kFamilyDenseSolve);
            if (band)
            {
                BL::toBand(BV, W.rhs, Z, s);
                BL::solve(BV, Z, D.st, s);
                BL::fromBand(BV, D.x, Z, true, s);
            }
            else
            {
                DL::toDense(W.rhs, n, npad, v, s);
                DL::solve(A, nt, v, D.st, s);
                DL::accumulate(D.x, v, n, true, s);
            }
        }
        check(hipEventRecord(ev[1], s), "event");
        DL::template residual<T>(P, W.S, W.rhs, D.x, D.r, D.rowNorm, D.norm, s);
        DL::store<T>(D.x, n, W.x, s);
        check(hipEventRecord(ev[2], s), "event");
    }

    // The outcome of denseEnqueue once the stream has been synchronised.
    DenseOutcome denseRead(const pb_ctx* ctx, double relTol)
    {
        const int32_t* hw = reinterpret_cast<const int32_t*>(ctx->pinned + kPinDense);
        const double* hn = ctx->pinned + kPinDense + 4;
        DenseOutcome out;
        if (hw[pbk::kDenseNotSpd]) return out;
        if (hw[pbk::kDenseTimeout]) throw std::runtime_error("dense triangular solve stopped waiting for a tile");
        out.factored = true;
        const double bn = std::sqrt(hn[1]);
        out.relResidual = bn > 0.0 ? std::sqrt(hn[0]) / bn : 0.0;
        out.converged = out.relResidual <= relTol;
        return out;
    }

    // The dense route of one LM step: FP64 Cholesky (fp64 work or a rejected seed), or with fp32Seed an FP32
    // factorisation refined in FP64 to relTol, accepted only when that residual is reached (else the step is
    // solved again with the FP64 factorisation). The solution goes to W.x.
    template <typename T>
    PcgResult denseSolve(pb_ctx* ctx, pbk::Work<T>& W, double relTol, bool fp32Seed, bool* seedAccepted,
                         pb_solve_report* rep)
    {
        const pbk::Problem& P = sysProblem(ctx);
        hipStream_t s = ctx->stream;
        PcgResult res;
        *seedAccepted = false;
        if (P.nUnits == 0)
        {
            res.converged = true;
            return res;
        }
        const int n = 6 * P.nUnits;
        DenseOutcome o;
        if (fp32Seed)
        {
            o = denseAttempt<float>(ctx, W, relTol, kDenseSeedRefinements, rep);
            *seedAccepted = o.factored && o.converged;
            res.iterations += o.refinements + (o.factored ? 1 : 0);
        }
        if (!*seedAccepted)
        {
            o = denseAttempt<double>(ctx, W, relTol, 0, rep);
            res.iterations += o.factored ? 1 : 0;
        }
        res.breakdown = !o.factored;
        res.converged = o.factored && std::isfinite(o.relResidual);
        res.rnorm = o.relResidual;
        if (res.converged) pbk::DenseLauncher<double>::store<T>(ctx->dense.x, n, W.x, s);
        return res;
    }

    struct Progress;

    // Reduced solve of one LM step by the route of the solve (dense Cholesky or PCG), with its report counters.
    template <typename T>
    PcgResult reducedSolve(pb_ctx* ctx, pbk::Work<T>& W, const pb_solve_options& o, Progress& pg,
                           pb_solve_report* rep, double relTol, double absTol, int maxPcg, int checkEvery,
                           double modelEta);

    void setMessage(pb_solve_report* r, const std::string& m)
    {
        std::snprintf(r->message, sizeof(r->message), "%s", m.c_str());
    }

    // Controller state of one solve, shared by the two parts of a mixed solve: when an FP32 step fails its line
    // search (path B) or is rejected (path A), the FP32 part hands the state, damping and counters to an FP64
    // continuation with the exact PCG stop of the FP64 path, which repeats the step from the same state (the FP32
    // rounding or the truncated PCG failed, not the state).
    struct Progress
    {
        bool started = false;
        bool inexact = false;     // truncated-Newton PCG stop and cap (a mixed solve, both parts)
        bool fallback = false;    // set by an FP32 part that hands over
        bool fp64Finish = false;  // the FP64 continuation of a mixed solve ran
        bool handOverBySwitch = false;  // the FP32 part handed over at its rounding floor (not after a failed step)
        int precision = PB_PRECISION_FP64;
        bool dense = false;       // reduced systems solved by the dense Cholesky route
        bool seed = false;        // dense route: FP32 factorisation refined in FP64 first (guarded)
        int fp32Accepted = 0;     // accepted steps of the FP32 part of a mixed solve
        int seedRejections = 0;   // consecutive FP32 seeds rejected (at kSeedRejectionsToStop the solve stops seeding)
        Clock::time_point t0{};
        double current = std::numeric_limits<double>::quiet_NaN();
        double damping = 0.0;
        int accepted = 0, rejected = 0, iterations = 0;
        bool converged = false, cancelled = false;
    };

    template <typename T>
    void runSolveAssembled(pb_ctx* ctx, const pb_solve_options& o, pb_solve_report* rep, const volatile int32_t* cancel,
                           Progress& pg);
    template <typename T>
    void publishState(pb_ctx* ctx, pb_solve_report* rep, const Progress& pg);

    template <typename T>
    PcgResult reducedSolve(pb_ctx* ctx, pbk::Work<T>& W, const pb_solve_options& o, Progress& pg,
                           pb_solve_report* rep, double relTol, double absTol, int maxPcg, int checkEvery,
                           double modelEta)
    {
        auto tl = Clock::now();
        PcgResult lin;
        double tol = relTol;
        if (pg.dense)
        {
            // an FP32-assembled system carries FP32 rounding itself, so its refinement stops at 1e-6; the
            // reference's mixed mode refines an FP32 seed to the FP64 tolerance of the GPU route (2e-10,
            // BundleAdjustPlaMatrix.cpp:326)
            const bool fp32Work = std::is_same<T, float>::value;
            tol = o.pcg_relative_tolerance > 0.0 ? o.pcg_relative_tolerance : (fp32Work ? 1e-6 : 2e-10);
            bool seed = false;
            lin = denseSolve(ctx, W, tol, pg.seed && pg.seedRejections < kSeedRejectionsToStop, &seed, rep);
            if (seed)
            {
                ++rep->mixed_seed_steps;
                pg.seedRejections = 0;
            }
            else if (pg.seed && pg.seedRejections < kSeedRejectionsToStop)
            {
                ++rep->dense_fallbacks;
                ++pg.seedRejections;
            }
        }
        else
        {
            lin = pcg(ctx, W, relTol, absTol, maxPcg, checkEvery, modelEta);
        }
        rep->linear_tolerance_minimum =
            rep->linear_tolerance_minimum > 0.0 ? std::min(rep->linear_tolerance_minimum, tol) : tol;
        rep->linear_tolerance_maximum = std::max(rep->linear_tolerance_maximum, tol);
        rep->linear_solve_seconds += seconds(tl);
        rep->linear_iterations += lin.iterations;
        return lin;
    }

    // SolverControlOptions::progressCallback after a trial step (BundleAdjustPlaMatrix.cpp:455-464,544-554):
    // false cancels the solve.
    bool reportProgress(pb_ctx* ctx, const pb_solve_options& o, const Progress& pg)
    {
        if (!ctx->progressFn) return true;
        pb_iteration_summary sm{};
        sm.current_iteration = pg.iterations;
        sm.max_iterations = std::max(1, o.max_iterations);
        sm.average_rms_pixels = std::sqrt(std::max(0.0, pg.current) / std::max(1, ctx->prep.observations));
        sm.valid_point_count = ctx->prep.activeTracks;
        return ctx->progressFn(&sm, ctx->progressUser) != 0;
    }

    // One LM step of the FP64 dense or band route (no FP32 seed) enqueued with one host wait: the assembly at
    // the damping, the reduced solve, back-substitution and the objective at the full step, each phase's results
    // copied to its own slot of the pinned area. The controller reads them in the order of the separate phases
    // and discards what follows a phase that failed; the kernels and their inputs are those of the separate
    // phases, so the step's numbers are the same. Phase times come from events between the phases.
    struct FusedStep
    {
        BuildResult build;
        DenseOutcome dense;
        double dd = 0.0, rmsG = 0.0, rmsP = 0.0, trialCost = 0.0;
        double assembly = 0.0, factor = 0.0, solve = 0.0, residual = 0.0, back = 0.0, trialStates = 0.0,
               objective = 0.0;
    };

    template <typename T>
    FusedStep fusedStep(pb_ctx* ctx, pbk::Work<T>& W, double damping, double relTol)
    {
        using Lr = pbk::Launcher<T>;
        const pbk::Problem& P = ctx->P;
        hipStream_t s = ctx->stream;
        hipEvent_t* ev = ctx->stepEv;
        double* h = ctx->pinned;
        check(hipEventRecord(ev[0], s), "event");
        buildLaunch(ctx, W, damping, true);
        check(hipEventRecord(ev[1], s), "event");
        denseEnqueue(ctx, W, ev + 2);
        {
            Span sp(ctx, 5);
            Lr::backSubstitute(P, W, s);
        }
        check(hipEventRecord(ev[5], s), "event");
        {
            Span sp(ctx, 4);
            check(hipEventRecord(ctx->trialEv[0], s), "event");
            Lr::trialStates(P, W, W.alphas, 1, s);
            check(hipEventRecord(ctx->trialEv[1], s), "event");
            Lr::trialObjective(P, W, W.alphas, 1, ctx->costBlocks, s);
        }
        // the build's objective and flags, the solver words and norms, the step's scalars and the unit step's
        // objective, straight to the pinned area (no kernel after the build writes the objective or the flags)
        Lr::publishStep(W, ctx->dense.st.words, ctx->dense.norm, h, kPinFlags, kPinDense, kPinStep, kPinTrial, s);
        check(hipEventRecord(ev[6], s), "event");
        check(hipStreamSynchronize(s), "step");
        FusedStep f;
        f.build = buildRead(ctx);
        f.dense = denseRead(ctx, relTol);
        f.dd = h[kPinStep + pbk::kScalarDd];
        f.rmsG = h[kPinStep + pbk::kScalarRmsGlobal];
        f.rmsP = h[kPinStep + pbk::kScalarRmsPoint];
        f.trialCost = h[kPinTrial];
        f.assembly = eventSeconds(ev[0], ev[1]);
        f.factor = eventSeconds(ev[1], ev[2]);
        f.solve = eventSeconds(ev[2], ev[3]);
        f.residual = eventSeconds(ev[3], ev[4]);
        f.back = eventSeconds(ev[4], ev[5]);
        f.trialStates = eventSeconds(ctx->trialEv[0], ctx->trialEv[1]);
        f.objective = eventSeconds(ev[5], ev[6]);
        return f;
    }

    // The reduced-solve outcome and report counters of a fused step (those of reducedSolve on the dense route
    // without a seed).
    PcgResult fusedLinear(const FusedStep& f, double tol, pb_solve_report* rep)
    {
        PcgResult lin;
        lin.iterations = f.dense.factored ? 1 : 0;
        lin.breakdown = !f.dense.factored;
        lin.converged = f.dense.factored && std::isfinite(f.dense.relResidual);
        lin.rnorm = f.dense.relResidual;
        rep->cholesky_factorization_seconds += f.factor;
        if (f.dense.factored)
        {
            rep->triangular_solve_seconds += f.solve;
            rep->residual_check_seconds += f.residual;
        }
        rep->linear_tolerance_minimum =
            rep->linear_tolerance_minimum > 0.0 ? std::min(rep->linear_tolerance_minimum, tol) : tol;
        rep->linear_tolerance_maximum = std::max(rep->linear_tolerance_maximum, tol);
        rep->linear_solve_seconds += f.factor + f.solve + f.residual;
        rep->linear_iterations += lin.iterations;
        return lin;
    }

    // The online-Schur controller (BundleAdjustPlaMatrix.cpp:162-631 with useReferenceOnlineSchur).
    template <typename T>
    void runSolve(pb_ctx* ctx, const pb_solve_options& o, pb_solve_report* rep, const volatile int32_t* cancel,
                  Progress& pg)
    {
        if (!onlineSchur(ctx, o))
        {
            runSolveAssembled<T>(ctx, o, rep, cancel, pg);
            return;
        }
        using Lr = pbk::Launcher<T>;
        pbk::Work<T>& W = work<T>(ctx);
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        const Prepared& pr = ctx->prep;
        const bool fp32 = std::is_same<T, float>::value;
        if (!pg.started)
        {
            resetState(ctx);
            check(hipStreamSynchronize(s), "reset");
            pg.t0 = Clock::now();
            pg.converged = pr.primaryBlocks == 0 && pr.trackBlocks == 0;
            pg.started = true;
        }
        const bool mixed = pg.inexact;
        const double relTol = o.pcg_relative_tolerance > 0.0 ? o.pcg_relative_tolerance : (mixed ? 1e-3 : 2e-10);
        const double absTol = o.pcg_absolute_tolerance > 0.0 ? o.pcg_absolute_tolerance : 1e-12;
        // iteration cap: the reference's max(200, 30 P) for exact solves; for truncated-Newton steps 0.6 P, which
        // leaves the aerial block 2.6x inside the 1e-6 final-cost bound (the cap trades PCG work for LM progress)
        const int maxPcg = o.pcg_max_iterations > 0 ? o.pcg_max_iterations
                                                    : (mixed ? std::max(200, 3 * pr.primaryBlocks / 5)
                                                             : std::max(200, 30 * pr.primaryBlocks));
        const int checkEvery = o.pcg_check_interval > 0 ? o.pcg_check_interval : 8;
        const int batch = o.armijo_batch > 0 ? std::min(o.armijo_batch, pbk::kMaxAlpha) : 4;
        const double modelEta = o.pcg_model_eta >= 0.0 ? o.pcg_model_eta : (mixed ? 0.1 : 0.0);
        const double rmsTol = o.step_rms_tolerance > 0.0 ? o.step_rms_tolerance : 1e-6;
        const double switchDecrease = o.mixed_switch_decrease == 0.0 ? kMixedSolveSwitchDecrease : o.mixed_switch_decrease;
        // the dense route's residual tolerance (as in reducedSolve)
        const double denseTol = o.pcg_relative_tolerance > 0.0 ? o.pcg_relative_tolerance : (fp32 ? 1e-6 : 2e-10);
        rep->linear_tolerance = relTol;
        ctx->lastScalarBytes = sizeof(T);
        rep->online_schur_used = 1;

        double& current = pg.current;
        double& damping = pg.damping;
        int &accepted = pg.accepted, &rejected = pg.rejected, &iterations = pg.iterations;
        bool& converged = pg.converged;
        bool& cancelled = pg.cancelled;
        auto isCancelled = [&]() { return cancelled || (cancel && *cancel != 0); };
        auto raise = [&]() { damping = damping == 0.0 ? 1e-3 : damping * 10.0; };
        if (converged && std::isnan(current))
        {
            const BuildResult b = build(ctx, W, 0.0);
            current = b.cost;
            rep->initial_cost = current;
            ++rep->objective_evaluations;
        }
        const int budget = std::max(1, o.max_iterations);
        bool terminate = false;
        double alphas[pbk::kMaxAlpha];
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
                auto ta = Clock::now();
                // the FP64 direct route without a seed takes the whole step with one host wait
                const bool fused = pg.dense && !(pg.seed && pg.seedRejections < kSeedRejectionsToStop) &&
                                   P.nUnits > 0 && o.line_search_steps > 0;
                FusedStep fs;
                BuildResult b;
                if (fused)
                {
                    fs = fusedStep(ctx, W, damping, denseTol);
                    b = fs.build;
                    rep->assembly_seconds += fs.assembly;
                }
                else
                {
                    b = build(ctx, W, damping);
                    rep->assembly_seconds += seconds(ta);
                }
                ++rep->linearizations;
                ++rep->objective_evaluations;
                current = b.cost;
                if (rep->linearizations == 1) rep->initial_cost = current;
                if (b.invalid) throw std::runtime_error("online Schur linearisation produced an invalid projection");
                if (b.singular || b.notSpd)
                {
                    // singular point block, or a reduced diagonal block that is not positive definite
                    // (a retryable "not positive definite" failure of the linear solve)
                    ++iterations;
                    ++rejected;
                    raise();
                    continue;
                }
                const PcgResult lin = fused ? fusedLinear(fs, denseTol, rep)
                                            : reducedSolve(ctx, W, o, pg, rep, relTol, absTol, maxPcg, checkEvery,
                                                           modelEta);
                if (lin.breakdown)
                {
                    ++iterations;
                    ++rejected;
                    raise();
                    continue;
                }
                ++iterations;
                if (!lin.converged)
                {
                    ++rejected;
                    raise();
                    continue;
                }
                double* h = ctx->pinned;
                double dd = fs.dd, rmsG = fs.rmsG, rmsP = fs.rmsP;
                if (fused)
                {
                    rep->back_substitution_seconds += fs.back;
                }
                else
                {
                    auto tb = Clock::now();
                    {
                        Span sp(ctx, 5);
                        Lr::backSubstitute(P, W, s);
                    }
                    check(hipMemcpyAsync(h, W.scalars, sizeof(double) * pbk::kScalarCount, hipMemcpyDeviceToHost, s),
                          "readback");
                    check(hipStreamSynchronize(s), "back-substitution");
                    rep->back_substitution_seconds += seconds(tb);
                    dd = h[pbk::kScalarDd];
                    rmsG = h[pbk::kScalarRmsGlobal];
                    rmsP = h[pbk::kScalarRmsPoint];
                }
                if (!std::isfinite(dd))
                {
                    ++rejected;
                    raise();
                    continue;
                }
                // Armijo backtracking; several step lengths per cost pass, decided in sequential order
                auto tobj = Clock::now();
                double alpha = 1.0, cand = kInf;
                bool armijo = false;
                int acceptedIndex = -1;
                int ls = 0;
                bool firstPass = true;
                while (ls < o.line_search_steps && !armijo)
                {
                    const int n = std::min(firstPass ? 1 : batch, o.line_search_steps - ls);
                    for (int k = 0; k < n; ++k) alphas[k] = alpha * std::ldexp(1.0, -k);
                    const double* costs = h;
                    if (fused && firstPass)
                    {
                        // the unit step's objective was evaluated with the step
                        costs = &fs.trialCost;
                        rep->trial_state_seconds += fs.trialStates;
                        rep->objective_seconds += fs.objective;
                    }
                    else
                    {
                        check(hipMemcpyAsync(W.alphas, alphas, sizeof(double) * n, hipMemcpyHostToDevice, s), "alphas");
                        {
                            Span sp(ctx, 4);
                            check(hipEventRecord(ctx->trialEv[0], s), "event");
                            Lr::trialStates(P, W, W.alphas, n, s);
                            check(hipEventRecord(ctx->trialEv[1], s), "event");
                            Lr::trialObjective(P, W, W.alphas, n, ctx->costBlocks, s);
                        }
                        check(hipMemcpyAsync(h, W.trialCost, sizeof(double) * n, hipMemcpyDeviceToHost, s), "readback");
                        check(hipStreamSynchronize(s), "trial cost");
                        rep->trial_state_seconds += eventSeconds(ctx->trialEv[0], ctx->trialEv[1]);
                    }
                    firstPass = false;
                    for (int k = 0; k < n; ++k)
                    {
                        ++rep->objective_evaluations;
                        ++ls;
                        cand = costs[k];
                        const double limit = current - o.armijo_coefficient * alphas[k] * dd;
                        if (std::isfinite(cand) && cand <= limit && cand < current)
                        {
                            armijo = true;
                            acceptedIndex = k;
                            alpha = alphas[k];
                            break;
                        }
                    }
                    if (!armijo) alpha = alphas[n - 1] * 0.5;
                }
                rep->objective_seconds += seconds(tobj);
                if (std::getenv("PB_TRACE"))
                    std::fprintf(stderr,
                                 "pb it=%d cost=%.12g cand=%.12g alpha=%g pcg=%d rnorm=%.3e rms_g=%.3e rms_p=%.3e "
                                 "dd=%.4e lambda=%g\n",
                                 iterations, current, cand, alpha, lin.iterations, lin.rnorm, rmsG, rmsP, dd, damping);
                if (armijo && std::isfinite(cand) && cand < current)
                {
                    Lr::accept(P, W, acceptedIndex, alpha, s);
                    const double before = current;
                    current = cand;
                    ++accepted;
                    if (fp32) ++pg.fp32Accepted;
                    damping /= 10.0;
                    converged = alpha * rmsG <= rmsTol && alpha * rmsP <= rmsTol;
                    relinearize = !converged;
                    // FP32 assembly has reached its rounding floor, or its step would stop the solve: the rest
                    // is assembled in FP64 (the stop is only taken on FP64 steps)
                    if (fp32 && switchDecrease > 0.0 && (converged || before - cand < switchDecrease * before))
                    {
                        converged = false;
                        pg.fallback = true;
                        pg.handOverBySwitch = true;
                        if (!reportProgress(ctx, o, pg)) cancelled = true;
                        return;
                    }
                }
                else
                {
                    ++rejected;
                    if (fp32)
                    {
                        pg.fallback = true;  // an FP64 continuation repeats the step from this state
                        return;
                    }
                    terminate = true;
                    relinearize = true;
                }
                if (!reportProgress(ctx, o, pg))
                {
                    cancelled = true;
                    break;
                }
            }
        }
        check(hipStreamSynchronize(s), "solve");
        publishState<T>(ctx, rep, pg);
    }

    // Final report fields and the refined state back to the host (both controllers).
    template <typename T>
    void publishState(pb_ctx* ctx, pb_solve_report* rep, const Progress& pg)
    {
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        const Prepared& pr = ctx->prep;
        const double current = pg.current;
        const bool converged = pg.converged, cancelled = pg.cancelled;
        rep->solve_seconds = seconds(pg.t0);
        // mixedPrecisionUsed (result.h:115-116): the reference mode counts accepted FP32 seeds; the FP32-assembly
        // mode counts FP32 steps the line search accepted (a solve handed to FP64 before any reports 0)
        rep->mixed_precision_used = pg.precision == PB_PRECISION_FP32_SEED ? (rep->mixed_seed_steps > 0 ? 1 : 0)
                                    : pg.precision == PB_PRECISION_MIXED   ? (pg.fp32Accepted > 0 ? 1 : 0)
                                                                           : 0;
        std::snprintf(rep->linear_solver_name, sizeof(rep->linear_solver_name), "%s",
                      !pg.dense                ? "pcg"
                      : ctx->band.active       ? (pg.seed ? "band-cr-fp32-refined" : "band-cr-fp64")
                      : pg.seed                ? "dense-cholesky-fp32-refined"
                                               : "dense-cholesky-fp64");
        if (pg.dense)
            std::snprintf(rep->preconditioner_name, sizeof(rep->preconditioner_name), "none");
        else if (ctx->cluster.active)
            std::snprintf(rep->preconditioner_name, sizeof(rep->preconditioner_name), "cluster-jacobi-%d",
                          ctx->cluster.size);
        else
            std::snprintf(rep->preconditioner_name, sizeof(rep->preconditioner_name), "block-jacobi-6");
        rep->accepted_steps = pg.accepted;
        rep->rejected_steps = pg.rejected;
        rep->iterations = pg.iterations;
        rep->final_cost = current;
        collectProfile(ctx);
        if (cancelled)
        {
            rep->solve_status = PB_SOLVE_CANCELLED;
            rep->solution_usable = 0;
            setMessage(rep, "cancelled; no intermediate solution published");
            ctx->haveResult = false;
            return;
        }
        rep->converged = converged ? 1 : 0;
        rep->solve_status = converged ? PB_SOLVE_SUCCESS : PB_SOLVE_NO_CONVERGENCE;
        rep->solution_usable = std::isfinite(current) ? 1 : 0;
        const std::string route = !pg.dense ? "block-Jacobi PCG" : ctx->band.active ? "band cyclic reduction"
                                                                                    : "dense Cholesky";
        std::string msg = converged ? (rep->online_schur_used ? "converged (online Schur + " + route + " + Armijo)"
                                                              : "converged (assembled Schur + " + route + " + line search)")
                                    : "iteration budget or line search exhausted";
        if (pg.fp64Finish)
            msg += pg.handOverBySwitch ? "; FP32 assembly handed over to FP64 at its rounding floor"
                                       : "; finished in FP64 with exact steps after a mixed step failed";
        setMessage(rep, msg);
        // refined state back to the host
        ctx->outR.resize(ctx->in.R.size());
        ctx->outC.resize(ctx->in.C.size());
        ctx->outGroupParams.resize(pr.gParams.size());
        ctx->outPointX.resize(ctx->lay.pointX.size());
        check(hipMemcpyAsync(ctx->outR.data(), P.camR, sizeof(double) * ctx->outR.size(), hipMemcpyDeviceToHost, s), "download");
        check(hipMemcpyAsync(ctx->outC.data(), P.camC, sizeof(double) * ctx->outC.size(), hipMemcpyDeviceToHost, s), "download");
        if (!ctx->outGroupParams.empty())
            check(hipMemcpyAsync(ctx->outGroupParams.data(), P.groupParams, sizeof(double) * ctx->outGroupParams.size(),
                                 hipMemcpyDeviceToHost, s),
                  "download");
        if (!ctx->outPointX.empty())
            check(hipMemcpyAsync(ctx->outPointX.data(), P.pointX, sizeof(double) * ctx->outPointX.size(),
                                 hipMemcpyDeviceToHost, s),
                  "download");
        if (ctx->rigSys)
        {
            const Input& in = ctx->in;
            ctx->outCapR.resize(in.capR.size());
            ctx->outCapC.resize(in.capC.size());
            ctx->outSenR.resize(in.senR.size());
            ctx->outSenC.resize(in.senC.size());
            check(hipMemcpyAsync(ctx->outCapR.data(), P.capR, sizeof(double) * in.capR.size(), hipMemcpyDeviceToHost, s), "download");
            check(hipMemcpyAsync(ctx->outCapC.data(), P.capC, sizeof(double) * in.capC.size(), hipMemcpyDeviceToHost, s), "download");
            check(hipMemcpyAsync(ctx->outSenR.data(), P.senR, sizeof(double) * in.senR.size(), hipMemcpyDeviceToHost, s), "download");
            check(hipMemcpyAsync(ctx->outSenC.data(), P.senC, sizeof(double) * in.senC.size(), hipMemcpyDeviceToHost, s), "download");
        }
        check(hipStreamSynchronize(s), "download");
        ctx->haveResult = true;
    }

    // The assembled-Schur controller (path A: BundleAdjustPlaMatrix.cpp:239-556 with the online Schur off).
    // One linearisation per outer step; a rejected step solves the same equations again with more damping (the
    // build is repeated at the unchanged state, which reproduces them); the directional derivative comes from an
    // objective probe at alpha = 1e-4 (:473-486) evaluated in the same cost pass as alpha = 1; any decrease is
    // accepted (:517-519), and a small step ends the solve without being applied (:451-465). A singular damped
    // point block is a linear-solver exception, which this path does not retry (:357-364); a reduced diagonal
    // block that is not positive definite counts as a solve that did not converge.
    template <typename T>
    void runSolveAssembled(pb_ctx* ctx, const pb_solve_options& o, pb_solve_report* rep, const volatile int32_t* cancel,
                           Progress& pg)
    {
        using Lr = pbk::Launcher<T>;
        pbk::Work<T>& W = work<T>(ctx);
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        const Prepared& pr = ctx->prep;
        const bool fp32 = std::is_same<T, float>::value;
        if (!pg.started)
        {
            resetState(ctx);
            check(hipStreamSynchronize(s), "reset");
            pg.t0 = Clock::now();
            pg.converged = pr.primaryBlocks == 0 && pr.trackBlocks == 0;
            pg.started = true;
        }
        const bool mixed = pg.inexact;
        const double relTol = o.pcg_relative_tolerance > 0.0 ? o.pcg_relative_tolerance : (mixed ? 1e-3 : 1e-12);
        const double absTol = o.pcg_absolute_tolerance > 0.0 ? o.pcg_absolute_tolerance : 1e-12;
        const int maxPcg = o.pcg_max_iterations > 0 ? o.pcg_max_iterations
                                                    : (mixed ? std::max(200, 3 * pr.primaryBlocks / 5)
                                                             : std::max(200, 30 * pr.primaryBlocks));
        const int checkEvery = o.pcg_check_interval > 0 ? o.pcg_check_interval : 8;
        const int batch = o.armijo_batch > 0 ? std::min(o.armijo_batch, pbk::kMaxAlpha) : 4;
        const double modelEta = o.pcg_model_eta >= 0.0 ? o.pcg_model_eta : (mixed ? 0.1 : 0.0);
        const double rmsTol = o.step_rms_tolerance > 0.0 ? o.step_rms_tolerance : 1e-6;
        const double switchDecrease = o.mixed_switch_decrease == 0.0 ? kMixedSolveSwitchDecrease : o.mixed_switch_decrease;
        constexpr double kProbe = 1e-4;
        rep->linear_tolerance = relTol;
        rep->online_schur_used = 0;
        ctx->lastScalarBytes = sizeof(T);

        double& current = pg.current;
        double& damping = pg.damping;
        int &accepted = pg.accepted, &rejected = pg.rejected, &iterations = pg.iterations;
        bool& converged = pg.converged;
        bool& cancelled = pg.cancelled;
        auto isCancelled = [&]() { return cancelled || (cancel && *cancel != 0); };
        auto raise = [&]() { damping = damping == 0.0 ? 1e-3 : damping * 10.0; };
        auto timedBuild = [&]() {
            auto ta = Clock::now();
            const BuildResult b = build(ctx, W, damping);
            rep->assembly_seconds += seconds(ta);
            if (b.invalid) throw std::runtime_error("assembled Schur linearisation produced an invalid projection");
            return b;
        };
        if (converged && std::isnan(current))
        {
            const BuildResult b = build(ctx, W, 0.0);
            current = b.cost;
            rep->initial_cost = current;
            ++rep->objective_evaluations;
        }
        const int budget = std::max(1, o.max_iterations);
        double alphas[pbk::kMaxAlpha];
        double* h = ctx->pinned;
        while (iterations < budget && !converged)
        {
            if (isCancelled())
            {
                cancelled = true;
                break;
            }
            BuildResult b = timedBuild();
            ++rep->linearizations;
            ++rep->objective_evaluations;
            current = b.cost;
            if (rep->linearizations == 1) rep->initial_cost = current;
            bool relinearize = false, fresh = true;
            while (iterations < budget && !converged && !relinearize)
            {
                if (isCancelled())
                {
                    cancelled = true;
                    break;
                }
                if (!fresh) b = timedBuild();  // same linearisation point, the current damping
                fresh = false;
                if (b.singular) throw std::runtime_error("linear solve failed: singular damped point block");
                PcgResult lin;
                if (!b.notSpd)
                    lin = reducedSolve(ctx, sysWork<T>(ctx), o, pg, rep, relTol, absTol, maxPcg, checkEvery, modelEta);
                ++iterations;
                if (b.notSpd || lin.breakdown || !lin.converged)
                {
                    ++rejected;
                    raise();
                    continue;
                }
                auto tb = Clock::now();
                {
                    Span sp(ctx, 5);
                    // a rig's step is solved over its captures and sensors; the cameras' steps follow from it
                    if (ctx->rigSys) Lr::rigExpand(P, sysWork<T>(ctx), W, s);
                    Lr::backSubstitute(P, W, s);
                    if (ctx->rigSys) Lr::rigStepRms(ctx->Psys, sysWork<T>(ctx), W, s);
                }
                check(hipMemcpyAsync(h, W.scalars, sizeof(double) * pbk::kScalarCount, hipMemcpyDeviceToHost, s),
                      "readback");
                check(hipStreamSynchronize(s), "back-substitution");
                rep->back_substitution_seconds += seconds(tb);
                const double rmsG = h[pbk::kScalarRmsGlobal], rmsP = h[pbk::kScalarRmsPoint];
                if (!std::isfinite(rmsG) || !std::isfinite(rmsP))
                    throw std::runtime_error("linear solve produced a non-finite step");
                if (rmsG <= rmsTol && rmsP <= rmsTol)
                {
                    converged = true;  // the small step is not applied
                    if (!reportProgress(ctx, o, pg)) cancelled = true;
                    continue;
                }
                // probe + line search: the first pass evaluates alpha = 1e-4 and 1, later passes `batch` halvings
                auto tobj = Clock::now();
                double alpha = 1.0, cand = kInf, candAlpha = 0.0, dd = 0.0;
                int candIndex = -1, ls = 0;
                bool armijo = false, probed = false;
                while (ls < o.line_search_steps && !armijo)
                {
                    int n = 0;
                    if (!probed) alphas[n++] = kProbe;
                    const int trials = std::min(probed ? batch : 1, o.line_search_steps - ls);
                    for (int k = 0; k < trials; ++k) alphas[n++] = alpha * std::ldexp(1.0, -k);
                    check(hipMemcpyAsync(W.alphas, alphas, sizeof(double) * n, hipMemcpyHostToDevice, s), "alphas");
                    {
                        Span sp(ctx, 4);
                        check(hipEventRecord(ctx->trialEv[0], s), "event");
                        Lr::trialStates(P, W, W.alphas, n, s);
                        check(hipEventRecord(ctx->trialEv[1], s), "event");
                        Lr::trialObjective(P, W, W.alphas, n, ctx->costBlocks, s);
                    }
                    check(hipMemcpyAsync(h, W.trialCost, sizeof(double) * n, hipMemcpyDeviceToHost, s), "readback");
                    check(hipStreamSynchronize(s), "trial cost");
                    rep->trial_state_seconds += eventSeconds(ctx->trialEv[0], ctx->trialEv[1]);
                    int k = 0;
                    if (!probed)
                    {
                        ++rep->objective_evaluations;
                        if (std::isfinite(h[0]) && h[0] < current) dd = (current - h[0]) / kProbe;
                        probed = true;
                        k = 1;
                    }
                    for (; k < n; ++k)
                    {
                        ++rep->objective_evaluations;
                        ++ls;
                        cand = h[k];
                        candIndex = k;
                        candAlpha = alphas[k];
                        const double limit = dd > 0.0 ? current - o.armijo_coefficient * alphas[k] * dd : current;
                        if (std::isfinite(cand) && cand <= limit && cand < current)
                        {
                            armijo = true;
                            alpha = alphas[k];
                            break;
                        }
                        alpha = alphas[k] * 0.5;
                    }
                }
                rep->objective_seconds += seconds(tobj);
                if (std::getenv("PB_TRACE"))
                    std::fprintf(stderr, "pb A it=%d cost=%.12g cand=%.12g alpha=%g rms_g=%.3e rms_p=%.3e dd=%.4e lambda=%g\n",
                                 iterations, current, cand, alpha, rmsG, rmsP, dd, damping);
                if (std::isfinite(cand) && cand < current)
                {
                    // the state of the last length tried; the stop test uses alpha, which is half of it when no
                    // length met the Armijo condition (:505-529)
                    Lr::accept(P, W, candIndex, candAlpha, s);
                    const double before = current;
                    current = cand;
                    ++accepted;
                    if (fp32) ++pg.fp32Accepted;
                    damping /= 10.0;
                    converged = alpha * rmsG <= rmsTol && alpha * rmsP <= rmsTol;
                    relinearize = !converged;
                    if (fp32 && switchDecrease > 0.0 && (converged || before - cand < switchDecrease * before))
                    {
                        converged = false;
                        pg.fallback = true;  // the rest is assembled in FP64 (as in runSolve)
                        pg.handOverBySwitch = true;
                        if (!reportProgress(ctx, o, pg)) cancelled = true;
                        return;
                    }
                }
                else
                {
                    ++rejected;
                    if (fp32)
                    {
                        pg.fallback = true;  // an FP64 continuation relinearises and repeats the step
                        return;
                    }
                    raise();
                }
                if (!reportProgress(ctx, o, pg))
                {
                    cancelled = true;
                    break;
                }
            }
        }
        check(hipStreamSynchronize(s), "solve");
        publishState<T>(ctx, rep, pg);
    }

    void fillReportCounts(const pb_ctx* ctx, pb_solve_report* r)
    {
        const Prepared& p = ctx->prep;
        r->observation_count = p.observations;
        r->active_track_count = p.activeTracks;
        r->rejected_initial_tracks = p.rejectedInitial;
        r->camera_block_count = p.refCameraBlocks();
        r->intrinsic_block_count = p.intrinsicBlocks;
        r->track_block_count = p.trackBlocks;
        r->primary_block_count = p.primaryBlocks;
        r->gauge_anchor_camera = p.anchor;
        r->gauge_scale_camera = p.scaleCam;
        r->gauge_baseline = p.baseline;
        std::snprintf(r->device_name, sizeof(r->device_name), "%s", ctx->deviceName.c_str());
    }

    pb_status fail(pb_ctx* ctx, pb_status st, const std::string& msg)
    {
        if (ctx) ctx->error = msg;
        return st;
    }

    template <typename F>
    pb_status guarded(pb_ctx* ctx, F&& f)
    {
        try
        {
            return f();
        }
        catch (const std::bad_alloc&)
        {
            return fail(ctx, PB_ERR_OOM, "out of memory");
        }
        catch (const HipError& e)
        {
            return fail(ctx, e.code == hipErrorOutOfMemory ? PB_ERR_OOM : PB_ERR_DEVICE, e.what());
        }
        catch (const std::invalid_argument& e)
        {
            const std::string m = e.what();
            return fail(ctx, m.rfind("unsupported", 0) == 0 ? PB_ERR_UNSUPPORTED : PB_ERR_INVALID, m);
        }
        catch (const std::exception& e)
        {
            return fail(ctx, PB_ERR_NUMERIC, e.what());
        }
        catch (...)
        {
            return fail(ctx, PB_ERR_NUMERIC, "unknown error");
        }
    }

    // Unit -> (reference 13-padded block, slot) mapping for dense exports.
    // Units are those of the reduced system the routes solve: a rig's units (rig blocks, then the group units) or
    // the camera-level ones.
    void unitSlots(const pb_ctx* ctx, int u, int* block, int slots[6])
    {
        const Prepared& p = ctx->prep;
        for (int a = 0; a < 6; ++a) slots[a] = -1;
        const int pose = p.refCameraBlocks();
        if (u < pose)
        {
            *block = u;
            for (int a = 0; a < 6; ++a) slots[a] = a;
            return;
        }
        const int cu = u - pose + p.cameraBlocks;  // the camera-level group unit
        for (int g = 0; g < p.G; ++g)
        {
            if (p.gK[g] == 0) continue;
            const int m = (p.gK[g] + 5) / 6;
            if (cu >= p.gUnit[g] && cu < p.gUnit[g] + m)
            {
                *block = pose + g;
                const int j = cu - p.gUnit[g];
                for (int a = 0; a < 6; ++a)
                    if (6 * j + a < p.gK[g]) slots[a] = p.gActive[pbdev::kMaxActive * g + 6 * j + a];
                return;
            }
        }
        *block = -1;
    }

    template <typename T>
    pb_status exportReduced(pb_ctx* ctx, const pb_solve_options& o, double damping, double* matrix, double* rhs,
                            double* cost, int32_t* primaryBlocks)
    {
        pbk::Work<T>& Wc = work<T>(ctx);
        resetState(ctx);
        const BuildResult b = build(ctx, Wc, damping);
        if (b.invalid) return fail(ctx, PB_ERR_NUMERIC, "invalid projection at the initial state");
        if (cost) *cost = b.cost;
        const int P13 = ctx->prep.primaryBlocks * 13;
        if (primaryBlocks) *primaryBlocks = ctx->prep.primaryBlocks;
        const pbk::Problem& P = sysProblem(ctx);
        const pbk::Work<T>& W = sysWork<T>(ctx);
        const Rows& rows = ctx->rigSys ? ctx->lay.sysRows : ctx->lay.rows;
        std::vector<T> S(size_t(P.nNnz) * 36);
        std::vector<double> r(size_t(P.nUnits) * 6);
        if (!S.empty())
            check(hipMemcpy(S.data(), W.S, sizeof(T) * S.size(), hipMemcpyDeviceToHost), "export");
        if (!r.empty()) check(hipMemcpy(r.data(), W.rhs, sizeof(double) * r.size(), hipMemcpyDeviceToHost), "export");
        if (matrix)
        {
            if (static_cast<long>(P13) * P13 > (1l << 26)) return fail(ctx, PB_ERR_UNSUPPORTED, "reduced system too large to export");
            std::fill(matrix, matrix + size_t(P13) * P13, 0.0);
            for (int i = 0; i < P13; ++i) matrix[size_t(i) * P13 + i] = 1.0;
            for (int u = 0; u < P.nUnits; ++u)
            {
                int bu, su[6];
                unitSlots(ctx, u, &bu, su);
                for (int k = rows.rowPtr[u]; k < rows.rowPtr[u + 1]; ++k)
                {
                    const int v = rows.colIdx[k];
                    int bv, sv[6];
                    unitSlots(ctx, v, &bv, sv);
                    for (int a = 0; a < 6; ++a)
                    {
                        if (su[a] < 0) continue;
                        for (int c = 0; c < 6; ++c)
                        {
                            if (sv[c] < 0) continue;
                            matrix[size_t(bu * 13 + su[a]) * P13 + bv * 13 + sv[c]] = double(S[size_t(k) * 36 + a * 6 + c]);
                        }
                    }
                }
            }
        }
        if (rhs)
        {
            std::fill(rhs, rhs + P13, 0.0);
            for (int u = 0; u < P.nUnits; ++u)
            {
                int bu, su[6];
                unitSlots(ctx, u, &bu, su);
                for (int a = 0; a < 6; ++a)
                    if (su[a] >= 0) rhs[bu * 13 + su[a]] = r[u * 6 + a];
            }
        }
        return PB_OK;
    }

    // ------------------------------------------------------------------------------------------
    // Batched windows (pb_hip_solve_batch).
    //
    // The windows are concatenated into one input (cameras, calibration groups and tracks numbered window by
    // window, observation cameras re-indexed) after each window's gauge is normalised on its own, and the
    // combined input is prepared and laid out as one problem. Its reduced system is block-diagonal by window,
    // and tracks, eliminated points, observations, pose units and group units are contiguous per window
    // (camera blocks and groups are numbered in camera order, tracks are ordered by their lowest camera).
    void prepareBatch(pb_ctx* ctx, const pb_batch_view* batch, const pb_solve_options& o)
    {
        releaseStructure(ctx);
        ctx->batchMode = false;
        ctx->uploaded = false;
        ctx->haveResult = false;
        const int nWin = batch->num_windows;
        BatchHost bh;
        bh.nWin = nWin;
        Input in;
        in.hasGroup = true;
        in.gaugePolicy = PB_GAUGE_CALLER_MANAGED;
        in.off.assign(1, 0);
        std::vector<int32_t> fixedCams, fixedTracks;
        bh.camBase.assign(1, 0);
        bh.trackBase.assign(1, 0);
        int nextGroup = 0;
        for (int w = 0; w < nWin; ++w)
        {
            const pb_problem_view* v = &batch->windows[w];
            auto windowError = [&](const std::string& m) {
                const bool unsupported = m.rfind("unsupported", 0) == 0;
                return std::invalid_argument((unsupported ? "unsupported: window " : "window ") + std::to_string(w) +
                                             ": " + m);
            };
            const std::string verr = v && (v->num_rig_captures > 0 || v->num_rig_sensors > 0)
                                         ? std::string("unsupported: camera rigs in batched windows")
                                         : validateInput(v);
            if (!verr.empty()) throw windowError(verr);
            if (v->num_control_points > 0 || v->num_laser_planes > 0)
                throw windowError("unsupported: survey constraints in a batched window");
            Input wi;
            copyIn(v, &wi);
            Prepared pw;
            const std::string gerr = normalizeGauge(wi, o, &pw);
            if (!gerr.empty()) throw windowError(gerr);
            const int cb = in.nCam, tb = in.nTracks;
            const int64_t ob = in.off.back();
            std::vector<std::pair<int32_t, int32_t>> ids;  // window group id -> combined id
            for (int c = 0; c < wi.nCam; ++c)
            {
                const int32_t id = wi.hasGroup ? wi.group[c] : 0;
                auto it = std::find_if(ids.begin(), ids.end(), [&](const auto& e) { return e.first == id; });
                if (it == ids.end())
                {
                    ids.emplace_back(id, nextGroup++);
                    in.group.push_back(ids.back().second);
                }
                else
                {
                    in.group.push_back(it->second);
                }
            }
            in.R.insert(in.R.end(), wi.R.begin(), wi.R.end());
            in.C.insert(in.C.end(), wi.C.begin(), wi.C.end());
            in.slots.insert(in.slots.end(), wi.slots.begin(), wi.slots.end());
            in.ref.insert(in.ref.end(), wi.ref.begin(), wi.ref.end());
            in.conv.insert(in.conv.end(), wi.conv.begin(), wi.conv.end());
            in.X.insert(in.X.end(), wi.X.begin(), wi.X.end());
            for (int t = 0; t < wi.nTracks; ++t) in.off.push_back(ob + wi.off[t + 1]);
            for (int32_t c : wi.obsCam) in.obsCam.push_back(c >= 0 && c < wi.nCam ? c + cb : -1);
            in.uv.insert(in.uv.end(), wi.uv.begin(), wi.uv.end());
            in.weight.insert(in.weight.end(), wi.weight.begin(), wi.weight.end());
            in.scale.insert(in.scale.end(), wi.scale.begin(), wi.scale.end());
            for (int32_t c : pw.fixedCams) fixedCams.push_back(c + cb);
            for (int32_t t : pw.fixedTracks) fixedTracks.push_back(t + tb);
            const bool gauge = pw.anchor >= 0 && pw.scaleCam >= 0 && pw.baseline > 0.0;
            bh.gaugeAnchor.push_back(gauge ? pw.anchor : -1);
            bh.gaugeScale.push_back(gauge ? pw.scaleCam : -1);
            bh.gaugeBaseline.push_back(gauge ? pw.baseline : 0.0);
            in.nCam += wi.nCam;
            in.nTracks += wi.nTracks;
            bh.camBase.push_back(in.nCam);
            bh.trackBase.push_back(in.nTracks);
        }
        ctx->in = std::move(in);
        Prepared p;
        p.fixedCams = fixedCams;
        p.fixedTracks = fixedTracks;  // ascending: window by window, each window's list sorted
        prepareActive(ctx->in, o, &p);
        prepareGroups(ctx->in, o, &p);
        Layout L;
        buildLayout(ctx->in, p, &L);
        uploadStructure(ctx, std::move(p), std::move(L), o);

        // window tables
        const Prepared& pp = ctx->prep;
        const Layout& LL = ctx->lay;
        const int nCam = ctx->in.nCam;
        auto windowOfTrack = [&](int t) {
            return static_cast<int32_t>(std::upper_bound(bh.trackBase.begin(), bh.trackBase.end(), t) -
                                        bh.trackBase.begin()) - 1;
        };
        auto offsetsOf = [&](const std::vector<int32_t>& win, int base, const char* what) {
            std::vector<int32_t> off(nWin + 1, 0);
            for (size_t i = 0; i < win.size(); ++i)
            {
                if (i > 0 && win[i] < win[i - 1])
                    throw std::logic_error(std::string("batch ") + what + " not grouped by window");
                ++off[win[i] + 1];
            }
            off[0] = base;
            for (int w = 0; w < nWin; ++w) off[w + 1] += off[w];
            return off;
        };
        std::vector<int32_t> camWin(nCam);
        for (int w = 0; w < nWin; ++w)
            for (int c = bh.camBase[w]; c < bh.camBase[w + 1]; ++c) camWin[c] = w;
        std::vector<int32_t> groupWin(pp.G, 0);
        for (int c = 0; c < nCam; ++c)
            if (pp.intrinsicBlock[c] >= 0) groupWin[pp.groupOfCam[c]] = camWin[c];
        std::vector<int32_t> elimWin(LL.nElim), obsWin(LL.nObs);
        for (int b = 0; b < LL.nElim; ++b) elimWin[b] = windowOfTrack(LL.activeTrackOf[LL.elimTrack[b]]);
        for (int i = 0; i < LL.nObs; ++i) obsWin[i] = windowOfTrack(LL.activeTrackOf[LL.obsPoint[i]]);
        std::vector<int32_t> unitWin(pp.nUnits, 0), poseWin, groupUnitWin;
        for (int c = 0; c < nCam; ++c)
            if (LL.camUnit[c] >= 0)
            {
                if (LL.camUnit[c] != static_cast<int32_t>(poseWin.size()))
                    throw std::logic_error("batch pose units not in camera order");
                unitWin[LL.camUnit[c]] = camWin[c];
                poseWin.push_back(camWin[c]);
            }
        for (int g = 0; g < pp.G; ++g)
        {
            if (pp.gK[g] == 0) continue;
            for (int j = 0; j < (pp.gK[g] + 5) / 6; ++j)
            {
                if (pp.gUnit[g] + j != pp.cameraBlocks + static_cast<int32_t>(groupUnitWin.size()))
                    throw std::logic_error("batch group units not in group order");
                unitWin[pp.gUnit[g] + j] = groupWin[g];
                groupUnitWin.push_back(groupWin[g]);
            }
        }
        const std::vector<int32_t> groupOffset = offsetsOf(groupWin, 0, "groups");
        const std::vector<int32_t> elimOffset = offsetsOf(elimWin, 0, "points");
        const std::vector<int32_t> obsOffset = offsetsOf(obsWin, 0, "observations");
        const std::vector<int32_t> poseUnitOffset = offsetsOf(poseWin, 0, "pose units");
        const std::vector<int32_t> groupUnitOffset = offsetsOf(groupUnitWin, pp.cameraBlocks, "group units");
        std::vector<int32_t> globalDim(nWin), chunkWin, chunkBegin, chunkEnd, winChunkOffset(1, 0);
        std::vector<int64_t> denseOffset(1, 0);
        int maxDim = 0;
        bh.observations.assign(nWin, 0);
        bh.activeTracks.assign(nWin, 0);
        bh.cameraBlocks.assign(nWin, 0);
        bh.primaryBlocks.assign(nWin, 0);
        for (int a = 0; a < LL.nActive; ++a) ++bh.activeTracks[windowOfTrack(LL.activeTrackOf[a])];
        for (int w = 0; w < nWin; ++w)
        {
            const int poseUnits = poseUnitOffset[w + 1] - poseUnitOffset[w];
            const int groupUnits = groupUnitOffset[w + 1] - groupUnitOffset[w];
            int dim = 6 * poseUnits;
            for (int g = groupOffset[w]; g < groupOffset[w + 1]; ++g) dim += pp.gK[g];
            globalDim[w] = dim;
            const int n = 6 * (poseUnits + groupUnits);
            if (n > pbk::kMaxDense)
                throw std::invalid_argument("unsupported: window " + std::to_string(w) + " has " +
                                            std::to_string(poseUnits + groupUnits) +
                                            " primary units; the batched dense solve takes at most " +
                                            std::to_string(pbk::kMaxDense / 6));
            maxDim = std::max(maxDim, n);
            const int64_t np = (n + pbk::kDenseTile - 1) / pbk::kDenseTile * pbk::kDenseTile;  // tile padding
            denseOffset.push_back(denseOffset.back() + np * np);
            for (int b = obsOffset[w]; b < obsOffset[w + 1]; b += pbk::kTrialChunk)
            {
                chunkWin.push_back(w);
                chunkBegin.push_back(b);
                chunkEnd.push_back(std::min(obsOffset[w + 1], b + pbk::kTrialChunk));
            }
            winChunkOffset.push_back(static_cast<int32_t>(chunkWin.size()));
            bh.observations[w] = obsOffset[w + 1] - obsOffset[w];
            bh.cameraBlocks[w] = poseUnits;
            bh.primaryBlocks[w] = poseUnits + (groupOffset[w + 1] - groupOffset[w]);
        }
        std::vector<int32_t> camGaugeAnchor(nCam, -1);
        for (int w = 0; w < nWin; ++w)
            if (bh.gaugeScale[w] >= 0)
                camGaugeAnchor[bh.camBase[w] + bh.gaugeScale[w]] = bh.camBase[w] + bh.gaugeAnchor[w];

        hipStream_t s = ctx->stream;
        DeviceAlloc& A = ctx->dbatch;
        pbk::Problem& P = ctx->P;
        pbk::Batch& B = bh.B;
        P.nWin = nWin;
        P.camWin = A.upload(camWin, s);
        P.groupWin = A.upload(groupWin, s);
        P.elimWin = A.upload(elimWin, s);
        P.unitWin = A.upload(unitWin, s);
        P.camGaugeAnchor = A.upload(camGaugeAnchor, s);
        P.winBaseline = A.upload(bh.gaugeBaseline, s);
        P.winFlags = A.alloc<int32_t>(size_t(nWin) * pbk::kFlagCount);
        {
            std::vector<int32_t> nnzRow(std::max(1, LL.nNnz));
            for (int u = 0; u < pp.nUnits; ++u)
                for (int k = LL.rows.rowPtr[u]; k < LL.rows.rowPtr[u + 1]; ++k) nnzRow[k] = u;
            P.nnzRow = A.upload(nnzRow, s);
        }
        B.nWin = nWin;
        B.nChunks = static_cast<int>(chunkWin.size());
        B.maxDim = maxDim;
        B.camOffset = A.upload(bh.camBase, s);
        B.groupOffset = A.upload(groupOffset, s);
        B.elimOffset = A.upload(elimOffset, s);
        B.poseUnitOffset = A.upload(poseUnitOffset, s);
        B.groupUnitOffset = A.upload(groupUnitOffset, s);
        B.globalDim = A.upload(globalDim, s);
        B.denseOffset = A.upload(denseOffset, s);
        B.chunkWin = A.upload(chunkWin, s);
        B.chunkBegin = A.upload(chunkBegin, s);
        B.chunkEnd = A.upload(chunkEnd, s);
        B.winChunkOffset = A.upload(winChunkOffset, s);
        B.status = A.alloc<int32_t>(nWin);
        B.active = A.alloc<int32_t>(nWin);
        B.reject = A.alloc<int32_t>(nWin);
        B.damping = A.alloc<double>(nWin);
        B.cost = A.alloc<double>(nWin);
        B.initialCost = A.alloc<double>(nWin);
        B.counters = A.alloc<int32_t>(size_t(nWin) * pbk::kWinCounterCount);
        B.dd = A.alloc<double>(nWin);
        B.rmsG = A.alloc<double>(nWin);
        B.rmsP = A.alloc<double>(nWin);
        B.searching = A.alloc<int32_t>(nWin);
        B.lsCount = A.alloc<int32_t>(nWin);
        B.alphaNext = A.alloc<double>(nWin);
        B.alpha = A.alloc<double>(size_t(nWin) * pbk::kMaxAlpha);
        B.nAlpha = A.alloc<int32_t>(nWin);
        B.acceptIdx = A.alloc<int32_t>(nWin);
        B.acceptAlpha = A.alloc<double>(nWin);
        B.acceptCost = A.alloc<double>(nWin);
        B.dense = A.alloc<double>(size_t(denseOffset.back()));
        B.trialPart = A.alloc<double>(size_t(B.nChunks) * pbk::kMaxAlpha);
        B.trialBad = A.alloc<int32_t>(size_t(B.nChunks) * pbk::kMaxAlpha);
        B.doneClock = A.alloc<unsigned long long>(nWin + 1);
        B.counts = A.alloc<int32_t>(4);
        B.phase = A.alloc<int32_t>(nWin);
        P.winDamping = B.damping;
        P.winActive = B.active;
        P.winAlpha = B.alpha;
        P.winNAlpha = B.nAlpha;
        int khz = 0;
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device), "hipDeviceGetAttribute");
        bh.clockHz = khz > 0 ? 1e3 * khz : 1e8;
        check(hipStreamSynchronize(s), "batch upload");
        ctx->bh = std::move(bh);
        ctx->uploaded = true;
        ctx->batchMode = true;
    }

    // One LM iteration of every running window of one precision phase (the kernels of that precision).
    // The three fixed launch sequences of a batch round for one assembly precision: the head (build, window
    // factorisations, back-substitution, the first trial pass of one step length), a further trial pass, and the
    // tail (controllers, state update, running-window counts). Each ends with the readback of the counters the host
    // polls. With graphs they are captured once per batch solve and replayed every round.
    struct BatchRoundGraphs
    {
        hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
        ~BatchRoundGraphs()
        {
            for (hipGraphExec_t& e : exec)
                if (e) (void)hipGraphExecDestroy(e);
        }
    };

    template <typename T>
    constexpr int precisionIndex()
    {
        return std::is_same<T, double>::value ? 1 : 0;
    }

    // shape < 0: every usable factorisation shape runs once, each between two events (the last one's result is
    // the one used, and all of them are the same)
    template <typename T>
    void batchSequence(pb_ctx* ctx, pbk::Work<T>& W, int phase, int passAlpha, int32_t* hc, int which, int shape)
    {
        using BL = pbk::BatchLauncher<T>;
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        const pbk::Batch& B = ctx->bh.B;
        if (which == 0)
        {
            BL::roundBegin(P, W, B, phase, s);
            BL::build(P, W, s);
            BL::afterBuild(P, W, B, s);
            if (shape >= 0)
                BL::factor(P, W, B, shape, s);
            else
            {
                FactorTuning& ft = ctx->tune;
                const int pi = precisionIndex<T>();
                for (int v = 0; v < pbk::kFactorShapes; ++v)
                {
                    if (!ft.usable[pi][v]) continue;
                    for (hipEvent_t& e : ft.ev[v])
                        if (!e)
                        {
                            check(hipEventCreate(&e), "hipEventCreate");
                            ctx->eventPool.push_back(e);  // destroyed with the context
                            ++ctx->eventNext;
                        }
                    check(hipEventRecord(ft.ev[v][0], s), "hipEventRecord");
                    BL::factor(P, W, B, v, s);
                    check(hipEventRecord(ft.ev[v][1], s), "hipEventRecord");
                }
            }
            BL::backSubstitute(P, W, B, s);
        }
        if (which <= 1)
        {
            BL::trialPass(P, W, B, which == 0 ? 1 : passAlpha, s);
            check(hipMemcpyAsync(hc + 2, B.counts + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s), "readback");
            return;
        }
        BL::finishRound(P, W, B, phase, s);
        BL::count(B, s);
        check(hipMemcpyAsync(hc, B.counts, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s), "readback");
    }

    template <typename T>
    void batchRun(pb_ctx* ctx, pbk::Work<T>& W, int phase, int passAlpha, int32_t* hc, int which,
                  BatchRoundGraphs* graphs, int shape)
    {
        hipStream_t s = ctx->stream;
        if (!graphs || shape < 0)
        {
            batchSequence<T>(ctx, W, phase, passAlpha, hc, which, shape);
            return;
        }
        if (!graphs->exec[which])
        {
            hipGraph_t g = nullptr;
            check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try
            {
                batchSequence<T>(ctx, W, phase, passAlpha, hc, which, shape);
            }
            catch (...)
            {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&graphs->exec[which], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            check(e, "hipGraphInstantiate");
        }
        check(hipGraphLaunch(graphs->exec[which], s), "hipGraphLaunch");
    }

    // Keeps the fastest factorisation shape of the timed head (ties and untimed shapes: the lower index).
    void pickFactorShape(pb_ctx* ctx, int pi)
    {
        FactorTuning& ft = ctx->tune;
        int best = -1;
        for (int v = 0; v < pbk::kFactorShapes; ++v)
        {
            ft.ms[pi][v] = -1.0;
            if (!ft.usable[pi][v]) continue;
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, ft.ev[v][0], ft.ev[v][1]), "hipEventElapsedTime");
            ft.ms[pi][v] = ms;
            if (best < 0 || ms < ft.ms[pi][best]) best = v;
        }
        ft.shape[pi] = best < 0 ? 0 : best;
    }

    template <typename T>
    int batchIteration(pb_ctx* ctx, pbk::Work<T>& W, int phase, const pb_solve_options& o, int passAlpha,
                       int32_t* hc, BatchRoundGraphs* graphs, int& shape)
    {
        hipStream_t s = ctx->stream;
        // trial passes until no window is still searching (each pass spends at least one of a window's
        // line-search steps, so at most line_search_steps passes); a head with shape < 0 times the factorisation
        // shapes and fixes the shape for the rest of the batch
        batchRun<T>(ctx, W, phase, passAlpha, hc, 0, graphs, shape);
        check(hipStreamSynchronize(s), "trial pass");
        if (shape < 0)
        {
            pickFactorShape(ctx, precisionIndex<T>());
            shape = ctx->tune.shape[precisionIndex<T>()];
        }
        int passes = 1;
        for (int pass = 1; pass <= o.line_search_steps && hc[2] != 0; ++pass)
        {
            batchRun<T>(ctx, W, phase, passAlpha, hc, 1, graphs, shape);
            check(hipStreamSynchronize(s), "trial pass");
            ++passes;
        }
        batchRun<T>(ctx, W, phase, passAlpha, hc, 2, graphs, shape);
        check(hipStreamSynchronize(s), "round");
        return passes;
    }

    template <typename T>
    void zeroDirect(pb_ctx* ctx, pbk::Work<T>& W)
    {
        const pbk::Problem& P = ctx->P;
        hipStream_t s = ctx->stream;
        check(hipMemsetAsync(W.D, 0, sizeof(double) * size_t(std::max(1, P.nBlocks)) * 36, s), "memset");
        check(hipMemsetAsync(W.directRhs, 0, sizeof(double) * size_t(std::max(1, P.nUnits)) * 6, s), "memset");
    }

    // One LM controller per window (BundleAdjustPlaMatrix.cpp:162-631 with useReferenceOnlineSchur, per window).
    // A round gives every running window one iteration: FP64 batches run the FP64 kernels; mixed batches run
    // the FP32 kernels on the windows still in the FP32 phase, then the FP64 kernels on the windows in the FP64
    // phase (a window that moves to FP64 in a round takes its first FP64 iteration in the same round).
    void runBatch(pb_ctx* ctx, const pb_solve_options& o, bool mixed, pb_window_report* reports,
                  pb_solve_report* summary, const volatile int32_t* cancel)
    {
        pbk::Work<double>& W64 = work<double>(ctx);
        pbk::Work<float>* W32 = mixed ? &work<float>(ctx) : nullptr;
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        BatchHost& bh = ctx->bh;
        pbk::Batch& B = bh.B;
        const int nWin = bh.nWin;
        resetState(ctx);
        B.maxIterations = std::max(1, o.max_iterations);
        B.lineSearchSteps = o.line_search_steps;
        B.armijo = o.armijo_coefficient;
        B.rmsTol = o.step_rms_tolerance > 0.0 ? o.step_rms_tolerance : 1e-6;
        B.switchDecrease = kMixedSwitchDecrease;
        const int passAlpha = o.armijo_batch > 0 ? std::min(o.armijo_batch, pbk::kMaxAlpha) : 4;
        // controller state (zero = running, no damping, no counts); direct blocks that no camera writes stay zero
        check(hipMemsetAsync(B.status, 0, sizeof(int32_t) * nWin, s), "memset");
        check(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(B.phase), mixed ? 0 : 1, nWin, s), "memset");
        check(hipMemsetAsync(B.damping, 0, sizeof(double) * nWin, s), "memset");
        check(hipMemsetAsync(B.counters, 0, sizeof(int32_t) * size_t(nWin) * pbk::kWinCounterCount, s), "memset");
        check(hipMemsetAsync(B.cost, 0, sizeof(double) * nWin, s), "memset");
        check(hipMemsetAsync(B.initialCost, 0, sizeof(double) * nWin, s), "memset");
        check(hipMemsetAsync(B.doneClock, 0, sizeof(unsigned long long) * (nWin + 1), s), "memset");
        zeroDirect(ctx, W64);
        if (W32) zeroDirect(ctx, *W32);
        FactorTuning& ft = ctx->tune;
        int winClass = 1;
        while (winClass < nWin) winClass *= 2;
        const std::vector<int32_t> key{winClass, (B.maxDim + 63) / 64};
        if (ft.key != key)
        {
            ft.key = key;
            ft.shape[0] = ft.shape[1] = -1;
        }
        pbk::BatchLauncher<double>::configure(B, ft.usable[1]);
        if (W32) pbk::BatchLauncher<float>::configure(B, ft.usable[0]);
        // shape per precision: fixed by the options, else the tuned one (-1: timed on this batch's first round)
        int shape[2];
        for (int pi = 0; pi < 2; ++pi)
        {
            shape[pi] = o.batch_factor_tiles       ? pbk::kFactorTiles
                        : o.batch_factor_shape >= 0 ? o.batch_factor_shape
                                                    : ft.shape[pi];
            if (!o.batch_factor_tiles && o.batch_factor_shape >= 0 && (pi == 1 || W32) &&
                !ft.usable[pi][o.batch_factor_shape])
                throw std::invalid_argument("unsupported: batch_factor_shape " + std::to_string(o.batch_factor_shape) +
                                  " needs more LDS than a workgroup has for windows of " +
                                  std::to_string(B.maxDim) + " unknowns");
        }
        check(hipStreamSynchronize(s), "reset");
        // captured once per batch solve (the launch arguments are this batch's), replayed every round
        BatchRoundGraphs graphs64, graphs32;
        BatchRoundGraphs* g64 = o.batch_graphs ? &graphs64 : nullptr;
        BatchRoundGraphs* g32 = o.batch_graphs ? &graphs32 : nullptr;
        const auto t0 = Clock::now();
        pbk::BatchLauncher<double>::start(B, s);
        int32_t* hc = reinterpret_cast<int32_t*>(ctx->pinned + kPinFlags);
        hc[0] = mixed ? nWin : 0;
        hc[1] = mixed ? 0 : nWin;
        int rounds = 0, passes = 0;
        bool cancelled = false;
        // every running window spends at least one iteration per round, so the budget ends the loop
        while (rounds <= B.maxIterations && hc[0] + hc[1] > 0)
        {
            if (cancel && *cancel)
            {
                cancelled = true;
                break;
            }
            if (W32 && hc[0] > 0) passes += batchIteration<float>(ctx, *W32, 0, o, passAlpha, hc, g32, shape[0]);
            if (hc[1] > 0) passes += batchIteration<double>(ctx, W64, 1, o, passAlpha, hc, g64, shape[1]);
            ++rounds;
        }
        check(hipStreamSynchronize(s), "batch");
        summary->solve_seconds = seconds(t0);
        summary->iterations = rounds;
        summary->objective_evaluations = passes;
        std::vector<int32_t> status(nWin), counters(size_t(nWin) * pbk::kWinCounterCount);
        std::vector<double> cost(nWin), initial(nWin);
        std::vector<unsigned long long> clk(nWin + 1);
        check(hipMemcpyAsync(status.data(), B.status, sizeof(int32_t) * nWin, hipMemcpyDeviceToHost, s), "readback");
        check(hipMemcpyAsync(counters.data(), B.counters, sizeof(int32_t) * counters.size(), hipMemcpyDeviceToHost, s),
              "readback");
        check(hipMemcpyAsync(cost.data(), B.cost, sizeof(double) * nWin, hipMemcpyDeviceToHost, s), "readback");
        check(hipMemcpyAsync(initial.data(), B.initialCost, sizeof(double) * nWin, hipMemcpyDeviceToHost, s),
              "readback");
        check(hipMemcpyAsync(clk.data(), B.doneClock, sizeof(unsigned long long) * (nWin + 1), hipMemcpyDeviceToHost, s),
              "readback");
        check(hipStreamSynchronize(s), "readback");
        int converged = 0;
        for (int w = 0; w < nWin; ++w)
        {
            pb_window_report& r = reports[w];
            const int32_t* cnt = &counters[size_t(w) * pbk::kWinCounterCount];
            r.iterations = cnt[pbk::kWinIterations];
            r.accepted_steps = cnt[pbk::kWinAccepted];
            r.rejected_steps = cnt[pbk::kWinRejected];
            r.linearizations = cnt[pbk::kWinLinearizations];
            r.objective_evaluations = cnt[pbk::kWinEvaluations];
            r.fp64_iterations = cnt[pbk::kWinFp64Iterations];
            r.observation_count = bh.observations[w];
            r.active_track_count = bh.activeTracks[w];
            r.camera_block_count = bh.cameraBlocks[w];
            r.primary_block_count = bh.primaryBlocks[w];
            r.initial_cost = initial[w];
            r.final_cost = cost[w];
            r.device_seconds = (clk[w] > 0 && clk[w] >= clk[nWin]) ? double(clk[w] - clk[nWin]) / bh.clockHz : 0.0;
            if (cancelled)
            {
                r.solve_status = PB_SOLVE_CANCELLED;
                continue;
            }
            if (bh.observations[w] == 0)
            {
                // as pb_hip_solve on the window alone: no usable tracks is invalid input
                r.solve_status = PB_SOLVE_INVALID_INPUT;
                continue;
            }
            r.converged = status[w] == pbk::kWinConverged ? 1 : 0;
            r.solve_status = status[w] == pbk::kWinConverged ? PB_SOLVE_SUCCESS
                             : status[w] == pbk::kWinNumeric ? PB_SOLVE_NUMERICAL_FAILURE
                                                             : PB_SOLVE_NO_CONVERGENCE;
            r.solution_usable = status[w] != pbk::kWinNumeric && std::isfinite(cost[w]) ? 1 : 0;
            converged += r.converged;
            summary->accepted_steps += r.accepted_steps;
            summary->rejected_steps += r.rejected_steps;
            summary->linearizations += r.linearizations;
            summary->initial_cost += r.initial_cost;
            summary->final_cost += r.final_cost;
        }
        if (cancelled)
        {
            summary->solve_status = PB_SOLVE_CANCELLED;
            setMessage(summary, "cancelled; no intermediate solution published");
            ctx->haveResult = false;
            return;
        }
        summary->converged = converged == nWin ? 1 : 0;
        summary->solution_usable = converged > 0 ? 1 : 0;
        summary->solve_status = converged == nWin ? PB_SOLVE_SUCCESS : PB_SOLVE_NO_CONVERGENCE;
        setMessage(summary, std::to_string(nWin) + " windows, " + std::to_string(converged) + " converged in " +
                                std::to_string(rounds) + " rounds (dense window solves + Armijo per window)");
        ctx->outR.resize(ctx->in.R.size());
        ctx->outC.resize(ctx->in.C.size());
        ctx->outGroupParams.resize(ctx->prep.gParams.size());
        ctx->outPointX.resize(ctx->lay.pointX.size());
        check(hipMemcpyAsync(ctx->outR.data(), P.camR, sizeof(double) * ctx->outR.size(), hipMemcpyDeviceToHost, s),
              "download");
        check(hipMemcpyAsync(ctx->outC.data(), P.camC, sizeof(double) * ctx->outC.size(), hipMemcpyDeviceToHost, s),
              "download");
        if (!ctx->outGroupParams.empty())
            check(hipMemcpyAsync(ctx->outGroupParams.data(), P.groupParams, sizeof(double) * ctx->outGroupParams.size(),
                                 hipMemcpyDeviceToHost, s),
                  "download");
        if (!ctx->outPointX.empty())
            check(hipMemcpyAsync(ctx->outPointX.data(), P.pointX, sizeof(double) * ctx->outPointX.size(),
                                 hipMemcpyDeviceToHost, s),
                  "download");
        check(hipStreamSynchronize(s), "download");
        ctx->haveResult = true;
    }

    // strictTrackRms (BundleAdjustQuality.cpp:302-363) of an input track at the input state, on the host (used for
    // the tracks that the solve did not activate).
    double hostStrictRms(const Input& in, const std::vector<pbdev::Camera>& cams, int t)
    {
        const double inf = std::numeric_limits<double>::infinity();
        const double* X = &in.X[3 * t];
        if (!std::isfinite(X[0]) || !std::isfinite(X[1]) || !std::isfinite(X[2])) return inf;
        double sum = 0.0;
        int residuals = 0, first = -1;
        bool second = false;
        for (int64_t i = in.off[t]; i < in.off[t + 1]; ++i)
        {
            const double sc = sanitizedScale(in.scale[i]);
            const double w = sanitizedWeight(in.weight[i]) / (sc * sc);
            if (!(w > 0.0)) continue;
            if (!usable(in, i)) return inf;
            const int c = in.obsCam[i];
            double u, v;
            if (!pbdev::project(cams[c], X[0], X[1], X[2], u, v)) return inf;
            const double du = u - in.uv[2 * i], dv = v - in.uv[2 * i + 1];
            sum += w * (du * du + dv * dv);
            residuals += 2;
            if (first < 0)
                first = c;
            else if (c != first)
                second = true;
        }
        if (residuals < 4 || !second) return inf;
        return std::sqrt(sum / residuals);
    }

    // Refined (or, before a solve, initial) state of input cameras [c0, c1) and tracks [t0, t1), written from
    // index 0 of each output (publishIntrinsics, BundleAdjustPlaMatrixModel.cpp:271-352, without transition
    // priors).
    void downloadRange(pb_ctx* ctx, int c0, int c1, int t0, int t1, double* R, double* C, double* intr, double* X,
                       int32_t* valid)
    {
        const Input& in = ctx->in;
        const bool res = ctx->haveResult;
        const Prepared& p = ctx->prep;
        for (int c = c0; c < c1; ++c)
        {
            const int k = c - c0;
            if (R) std::copy_n(res ? &ctx->outR[9 * c] : &in.R[9 * c], 9, R + 9 * k);
            if (C) std::copy_n(res ? &ctx->outC[3 * c] : &in.C[3 * c], 3, C + 3 * k);
            if (intr)
            {
                double pub[13];
                std::copy_n(&in.slots[13 * c], 13, pub);
                if (res && p.G > 0 && p.intrinsicBlock[c] >= 0)
                {
                    const int g = p.groupOfCam[c];
                    const uint32_t en = p.gEnabled[g];
                    const double* gp = &ctx->outGroupParams[13 * g];
                    const double* ref = &in.ref[13 * c];
                    pub[0] = (en & 1u) ? gp[0] : in.slots[13 * c];
                    pub[1] = (en & 2u) ? gp[1] : in.slots[13 * c + 1];
                    pub[2] = (en & 4u) ? ref[2] + gp[2] : in.slots[13 * c + 2];
                    pub[3] = (en & 8u) ? ref[3] + gp[3] : in.slots[13 * c + 3];
                    for (int q = 4; q < 13; ++q)
                        if ((en >> q) & 1u) pub[q] = gp[q];
                }
                std::copy_n(pub, 13, intr + 13 * k);
            }
        }
        for (int t = t0; t < t1; ++t)
        {
            const int a = ctx->structureValid ? ctx->lay.activeIndexOf[t] : -1;
            const int k = t - t0;
            if (X) std::copy_n((res && a >= 0) ? &ctx->outPointX[3 * a] : &in.X[3 * t], 3, X + 3 * k);
            if (valid) valid[k] = (ctx->structureValid && p.activeTrack[t]) ? 1 : 0;
        }
    }
}  // namespace

extern "C" {

void pb_default_options(pb_solve_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iterations = 20;
    o->robust_loss = PB_LOSS_LEAST_SQUARES;
    o->robust_loss_scale = 1.0;
    o->armijo_coefficient = 1e-3;
    o->line_search_steps = 34;
    o->use_online_schur = 0;  // useReferenceOnlineSchur = false (options.h:65)
    o->max_initial_track_rms = 100.0;
    o->laser_plane_weight = 1.0;
    o->laser_huber_delta = 0.2;
    o->control_point_weight = 1.0;
    o->control_point_huber_delta = 0.2;
    o->use_transition_prior = 0;
    o->previous_intrinsic_mask = 0;
    o->batch_graphs = 1;
    o->batch_factor_shape = -1;
    o->batch_factor_tiles = 1;
    o->pcg_graphs = 1;
    o->refine_camera_pose = 1;
    o->shared_intrinsic_mask = 0;
    o->min_focal_scale = 0.5;
    o->max_focal_scale = 4.0;
    o->min_aspect_scale = 0.85;
    o->max_aspect_scale = 1.18;
    o->max_principal_offset_fraction = 0.08;
    o->focal_prior_sigma = 0.35;
    o->principal_prior_sigma_fraction = 0.04;
    o->aspect_prior_sigma = 0.08;
    const double mx[9] = {0.35, 0.35, 0.35, 0.02, 0.02, 0.10, 0.35, 0.35, 0.35};
    const double sg[9] = {0.15, 0.15, 0.15, 0.01, 0.01, 0.02, 0.15, 0.15, 0.15};
    for (int i = 0; i < 9; ++i)
    {
        o->max_abs[i] = mx[i];
        o->prior_sigma[i] = sg[i];
    }
    o->low_order_distortion_scale = 1.0;
    o->device = 0;
    o->precision = PB_PRECISION_FP64;
    o->pcg_model_eta = -1.0;
    o->reduced_solver = PB_REDUCED_AUTO;
    o->dense_max_rows = 0;
    o->preconditioner_cluster_size = 1;
}

int pb_hip_available(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

pb_status pb_hip_create(int device, pb_ctx** out)
{
    if (!out) return PB_ERR_INVALID;
    *out = nullptr;
    pb_ctx* ctx = nullptr;
    try
    {
        ctx = new pb_ctx();
    }
    catch (const std::bad_alloc&)
    {
        return PB_ERR_OOM;
    }
    const pb_status st = guarded(ctx, [&]() -> pb_status {
        if (!pb_hip_available(device)) return fail(ctx, PB_ERR_DEVICE, "no gfx950 device at this index");
        ctx->device = device;
        check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        // the marketing name can be empty on some systems; the architecture name and the compute-unit count
        // identify the device then
        const std::string marketing = prop.name;
        ctx->deviceName = (marketing.empty() ? std::string("AMD Instinct (") + std::to_string(prop.multiProcessorCount) +
                                                   " CUs)"
                                             : marketing) +
                          " (" + prop.gcnArchName + ")";
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (hipEvent_t& e : ctx->trialEv) check(hipEventCreate(&e), "hipEventCreate");
        for (hipEvent_t& e : ctx->stepEv) check(hipEventCreate(&e), "hipEventCreate");
        void* p = nullptr;
        check(hipHostMalloc(&p, 4096, hipHostMallocCoherent), "hipHostMalloc");  // kernels publish step results here
        ctx->pinned = static_cast<double*>(p);
        return PB_OK;
    });
    if (st != PB_OK)
    {
        pb_hip_destroy(ctx);
        return st;
    }
    *out = ctx;
    return PB_OK;
}

pb_status pb_hip_destroy(pb_ctx* ctx)
{
    if (!ctx) return PB_OK;
    if (ctx->stream)
    {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    ctx->dstruct.release();
    ctx->dwork32.release();
    ctx->dwork64.release();
    ctx->dbatch.release();
    ctx->dquality.release();
    for (hipEvent_t e : ctx->eventPool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->trialEv)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->stepEv)
        if (e) (void)hipEventDestroy(e);
    ctx->band.tables.release();
    ctx->band.arena32.release();
    ctx->band.arena64.release();
    for (auto& g : ctx->pcgGraph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->dense.vec.release();
    ctx->dense.mat32.release();
    ctx->dense.mat64.release();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PB_OK;
}

pb_status pb_hip_upload(pb_ctx* ctx, const pb_problem_view* problem)
{
    if (!ctx) return PB_ERR_INVALID;
    const std::string err = validateInput(problem);
    if (!err.empty()) return fail(ctx, PB_ERR_INVALID, err);
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->uploaded = false;
        ctx->structureValid = false;
        ctx->haveResult = false;
        copyIn(problem, &ctx->in);
        if (ctx->batchMode) releaseStructure(ctx);
        ctx->batchMode = false;
        ctx->structureValid = false;
        ctx->haveResult = false;
        ctx->uploaded = true;
        return PB_OK;
    });
}

pb_status pb_hip_solve(pb_ctx* ctx, const pb_solve_options* options, pb_solve_report* report,
                       const volatile int32_t* cancel)
{
    if (!ctx || !report) return PB_ERR_INVALID;
    std::memset(report, 0, sizeof(*report));
    report->gauge_anchor_camera = report->gauge_scale_camera = -1;
    const std::string oerr = validateOptions(options);
    if (!oerr.empty())
    {
        report->solve_status = PB_SOLVE_INVALID_INPUT;
        setMessage(report, oerr);
        return fail(ctx, PB_ERR_INVALID, oerr);
    }
    if (!ctx->uploaded || ctx->batchMode)
    {
        report->solve_status = PB_SOLVE_INVALID_INPUT;
        setMessage(report, "no problem uploaded");
        return fail(ctx, PB_ERR_INVALID, "no problem uploaded");
    }
    // BundleAdjustValidation.cpp:589-595
    if (ctx->in.nPlanes > 0 && (!std::isfinite(options->laser_plane_weight) || options->laser_plane_weight <= 0.0 ||
                                !std::isfinite(options->laser_huber_delta) || options->laser_huber_delta < 0.0))
    {
        report->solve_status = PB_SOLVE_INVALID_INPUT;
        setMessage(report, "laser-plane weight must be positive and its Huber threshold finite and non-negative");
        return fail(ctx, PB_ERR_INVALID, report->message);
    }
    ctx->profile = options->profile_kernels != 0;
    const pb_status st = guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (cancel && *cancel)
        {
            report->solve_status = PB_SOLVE_CANCELLED;
            setMessage(report, "cancelled before start");
            ctx->haveResult = false;
            return PB_ERR_CANCELLED;
        }
        const auto ts = Clock::now();
        const int64_t builds = ctx->patternBuilds;
        ensureStructure(ctx, *options);
        fillReportCounts(ctx, report);
        report->schur_pattern_builds = int32_t(ctx->patternBuilds - builds);
        report->schur_pattern_reuses = report->schur_pattern_builds ? 0 : 1;
        report->schur_assembly_on_device = 1;
        applySolveOptions(ctx, *options);
        if (ctx->in.nCam == 0 || ctx->in.nTracks == 0)
        {
            report->solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(report, "empty cameras or tracks");
            return PB_ERR_INVALID;
        }
        if (ctx->prep.observations == 0)
        {
            report->solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(report, "no usable tracks");
            return PB_ERR_INVALID;
        }
        if (options->precision == PB_PRECISION_MIXED)
            (void)work<float>(ctx);
        else
            (void)work<double>(ctx);
        Progress pg;
        pg.precision = options->precision;
        const Route route = reducedRoute(ctx, *options);
        pg.dense = route != kRoutePcg;
        ctx->band.active = route == kRouteBand;
        // the FP32-seed mode always seeds; the mixed mode factors its FP32 system in FP32 (refined to 1e-6) on the
        // dense route, where that halves the MFMA time, and in FP64 on the band route, whose FP32 chain is only a
        // quarter shorter and would add refinement solves and host waits to every step
        pg.seed = pg.dense && (options->precision == PB_PRECISION_FP32_SEED ||
                               (options->precision == PB_PRECISION_MIXED && !ctx->band.active));
        if (pg.dense) ensureDense(ctx, denseTiles(ctx->P), pg.seed, !ctx->band.active);
        if (ctx->band.active)
        {
            (void)bandArena<double>(ctx);
            if (options->precision != PB_PRECISION_FP64) (void)bandArena<float>(ctx);
        }
        // the PCG route's preconditioner: per-unit blocks in the fused iteration, or clusters of primary blocks
        ctx->cluster.active = false;
        if (!pg.dense) planClusters(ctx, options->preconditioner_cluster_size);
        if (pg.dense) ctx->cluster.active = false;
        check(hipStreamSynchronize(ctx->stream), "setup");
        report->setup_seconds = seconds(ts);
        ctx->lastOptions = *options;
        // the truncated-Newton PCG stop is a property of the FP32-assembly mode on the PCG route
        pg.inexact = options->precision == PB_PRECISION_MIXED && !pg.dense;
        if (options->precision == PB_PRECISION_MIXED)
        {
            runSolve<float>(ctx, *options, report, cancel, pg);
            if (pg.fallback)
            {
                (void)work<double>(ctx);
                pg.fallback = false;
                pg.fp64Finish = true;
                pg.inexact = false;
                pg.seed = false;
                runSolve<double>(ctx, *options, report, cancel, pg);
            }
        }
        else
        {
            runSolve<double>(ctx, *options, report, cancel, pg);
        }
        ctx->lastSolveStatus = report->solve_status;
        ctx->lastUsable = report->solution_usable;
        if (report->solve_status == PB_SOLVE_CANCELLED) return PB_ERR_CANCELLED;
        return PB_OK;
    });
    if (st != PB_OK && st != PB_ERR_CANCELLED && report->solve_status == PB_SOLVE_NOT_RUN)
    {
        report->solve_status = st == PB_ERR_UNSUPPORTED  ? PB_SOLVE_UNSUPPORTED
                               : st == PB_ERR_INVALID    ? PB_SOLVE_INVALID_INPUT
                               : st == PB_ERR_DEVICE     ? PB_SOLVE_BACKEND_UNAVAILABLE
                                                         : PB_SOLVE_NUMERICAL_FAILURE;
        setMessage(report, ctx->error);
    }
    if (st == PB_ERR_NUMERIC)
    {
        report->solve_status = PB_SOLVE_NUMERICAL_FAILURE;
        setMessage(report, "numerical failure: " + ctx->error);
        ctx->haveResult = false;
    }
    return st;
}

pb_status pb_hip_backward(pb_ctx* ctx, const double* dl_dcenter, const double* dl_dpoint, double* dl_dobs,
                          double* dl_dcontrol)
{
    return pb_hip_backward_pose(ctx, dl_dcenter, nullptr, dl_dpoint, dl_dobs, dl_dcontrol);
}

pb_status pb_hip_backward_pose(pb_ctx* ctx, const double* dl_dcenter, const double* dl_drotation,
                               const double* dl_dpoint, double* dl_dobs, double* dl_dcontrol)
{
    if (!ctx || !dl_dobs) return PB_ERR_INVALID;
    if (!ctx->uploaded || !ctx->haveResult || ctx->batchMode)
        return fail(ctx, PB_ERR_INVALID, "pb_hip_backward needs the result of a pb_hip_solve on this context");
    if (ctx->in.rig()) return fail(ctx, PB_ERR_UNSUPPORTED, "unsupported: the adjoint of a camera rig problem");
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        using Lr = pbk::Launcher<double>;
        pbk::Work<double>& W = work<double>(ctx);
        pbk::Problem& P = ctx->P;
        const Layout& L = ctx->lay;
        hipStream_t s = ctx->stream;
        applySolveOptions(ctx, ctx->lastOptions);
        // loss gradients in device order (cameras as given, points in elimination order)
        std::vector<double> hc(3 * size_t(std::max(1, P.nCam)), 0.0), hp(3 * size_t(std::max(1, P.nElim)), 0.0);
        std::vector<double> hr(9 * size_t(std::max(1, P.nCam)), 0.0);
        if (dl_dcenter) std::copy_n(dl_dcenter, 3 * size_t(P.nCam), hc.begin());
        if (dl_drotation) std::copy_n(dl_drotation, 9 * size_t(P.nCam), hr.begin());
        if (dl_dpoint)
            for (int b = 0; b < P.nElim; ++b)
                std::copy_n(dl_dpoint + 3 * size_t(L.activeTrackOf[L.elimTrack[b]]), 3, hp.begin() + 3 * b);
        DeviceAlloc tmp;
        const double* dC = tmp.upload(hc, s);
        const double* dR = tmp.upload(hr, s);
        const double* dP = tmp.upload(hp, s);
        double* dOut = tmp.alloc<double>(2 * size_t(std::max(1, P.nObs)));
        double* dCtl = tmp.alloc<double>(3 * size_t(std::max(1, ctx->in.nControl)));
        check(hipMemsetAsync(dCtl, 0, sizeof(double) * 3 * size_t(std::max(1, ctx->in.nControl)), s), "memset");
        // the normal equations at the converged state with the adjoint weight roots, no damping
        P.adjoint = 1;
        BuildResult br;
        try
        {
            br = build(ctx, W, 0.0);
        }
        catch (...)
        {
            P.adjoint = 0;
            throw;
        }
        P.adjoint = 0;
        const long nRows = 6L * P.nUnits;
        // the route of the last solve when it was forced to the dense Cholesky, else band when it fits
        ctx->band.active = ctx->lastOptions.reduced_solver != PB_REDUCED_DENSE && planBand(ctx);
        const bool direct = ctx->band.active || (nRows > 0 && nRows <= kDenseAdjointMax);
        // a 6x6 diagonal block that is not positive definite only stops the PCG route; the direct routes report
        // through their factorisation
        if (br.invalid || br.singular || (br.notSpd && !direct))
        {
            tmp.release();
            return fail(ctx, PB_ERR_NUMERIC, br.invalid    ? "adjoint linearisation has an invalid projection"
                                             : br.singular ? "adjoint system has a singular point block"
                                                           : "adjoint system has a diagonal block that is not positive definite");
        }
        check(hipMemsetAsync(W.directRhs, 0, sizeof(double) * size_t(std::max(1, P.nUnits)) * 6, s), "memset");
        Lr::adjointRhs(P, W, dC, dR, dP, s);
        Lr::schurBlocks(P, W, 0.0, s);
        // H lambda = g is solved once per call, exactly: the dense FP64 Cholesky of the reduced system up to
        // kDenseAdjointMax rows (the aerial block has 6,024; its condition number near 1e14 keeps PCG from a
        // tight residual; 60k rows of a 10k-camera block take 29 GB of HBM), FP64 PCG to a 1e-10 relative
        // residual above that
        if (direct)
        {
            bool seed = false;
            const PcgResult lin = denseSolve(ctx, W, 0.0, false, &seed, nullptr);
            if (lin.breakdown)
            {
                tmp.release();
                return fail(ctx, PB_ERR_NUMERIC, "adjoint system is not positive definite");
            }
        }
        else
        {
            const PcgResult lin = pcg(ctx, W, 1e-10, 1e-300, std::max(200, 30 * ctx->prep.primaryBlocks), 8, 0.0);
            if (!lin.converged)
            {
                tmp.release();
                return fail(ctx, PB_ERR_NUMERIC, "adjoint PCG did not converge");
            }
        }
        Lr::backSubstitute(P, W, s);
        Lr::adjointObservations(P, W, dOut, s);
        Lr::adjointConstraints(P, W, dCtl, s);
        std::vector<double> out(2 * size_t(P.nObs));
        if (P.nObs > 0)
            check(hipMemcpyAsync(out.data(), dOut, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s), "readback");
        if (dl_dcontrol && ctx->in.nControl > 0)
            check(hipMemcpyAsync(dl_dcontrol, dCtl, sizeof(double) * 3 * size_t(ctx->in.nControl), hipMemcpyDeviceToHost, s),
                  "readback");
        check(hipStreamSynchronize(s), "adjoint");
        tmp.release();
        const int64_t nIn = ctx->in.off.empty() ? 0 : ctx->in.off.back();
        std::fill_n(dl_dobs, 2 * size_t(nIn), 0.0);
        for (int o = 0; o < P.nObs; ++o)
        {
            dl_dobs[2 * L.obsInput[o]] = out[2 * o];
            dl_dobs[2 * L.obsInput[o] + 1] = out[2 * o + 1];
        }
        return PB_OK;
    });
}

pb_status pb_hip_get_rig_masks(pb_ctx* ctx, int32_t* captureBlock, int32_t* sensorBlock)
{
    if (!ctx || !ctx->structureValid) return fail(ctx, PB_ERR_INVALID, "no prepared problem (solve first)");
    const Prepared& p = ctx->prep;
    for (int c = 0; c < ctx->in.nCam; ++c)
    {
        if (captureBlock) captureBlock[c] = p.rig ? p.rigCapBlockByCam[c] : -1;
        if (sensorBlock) sensorBlock[c] = p.rig ? p.rigSenBlockByCam[c] : -1;
    }
    return PB_OK;
}

pb_status pb_hip_download_rig(pb_ctx* ctx, double* capR, double* capC, double* senR, double* senC)
{
    if (!ctx || !ctx->uploaded) return fail(ctx, PB_ERR_INVALID, "no uploaded problem");
    const Input& in = ctx->in;
    const bool refined = ctx->haveResult && ctx->rigSys && !ctx->batchMode;
    auto put = [](double* dst, const std::vector<double>& src) {
        if (dst && !src.empty()) std::copy(src.begin(), src.end(), dst);
    };
    put(capR, refined ? ctx->outCapR : in.capR);
    put(capC, refined ? ctx->outCapC : in.capC);
    put(senR, refined ? ctx->outSenR : in.senR);
    put(senC, refined ? ctx->outSenC : in.senC);
    return PB_OK;
}

pb_status pb_hip_download(pb_ctx* ctx, double* R, double* C, double* intr, double* X, int32_t* valid)
{
    if (!ctx || !ctx->uploaded) return PB_ERR_INVALID;
    downloadRange(ctx, 0, ctx->in.nCam, 0, ctx->in.nTracks, R, C, intr, X, valid);
    return PB_OK;
}

pb_status pb_hip_batch_download(pb_ctx* ctx, int32_t window, double* R, double* C, double* intr, double* X,
                                int32_t* valid)
{
    if (!ctx) return PB_ERR_INVALID;
    if (!ctx->batchMode) return fail(ctx, PB_ERR_INVALID, "the context holds no batch");
    if (window < 0 || window >= ctx->bh.nWin) return fail(ctx, PB_ERR_INVALID, "window index out of range");
    const BatchHost& bh = ctx->bh;
    downloadRange(ctx, bh.camBase[window], bh.camBase[window + 1], bh.trackBase[window], bh.trackBase[window + 1], R,
                  C, intr, X, valid);
    return PB_OK;
}

static_assert(PB_BATCH_FACTOR_SHAPES == pbk::kFactorShapes, "factorisation shapes of the ABI and the kernels");

pb_status pb_hip_batch_factor_shape(int32_t shape, int32_t* threads, int32_t* panel)
{
    if (shape < 0 || shape >= pbk::kFactorShapes) return PB_ERR_INVALID;
    if (threads) *threads = pbk::kFactorShape[shape].threads;
    if (panel) *panel = pbk::kFactorShape[shape].panel;
    return PB_OK;
}

pb_status pb_hip_batch_tuning(pb_ctx* ctx, int32_t fp64, int32_t* shape, double* ms)
{
    if (!ctx || !shape || (fp64 != 0 && fp64 != 1)) return PB_ERR_INVALID;
    const FactorTuning& ft = ctx->tune;
    *shape = ft.shape[fp64];
    if (ms)
        for (int v = 0; v < pbk::kFactorShapes; ++v) ms[v] = ft.shape[fp64] >= 0 ? ft.ms[fp64][v] : -1.0;
    return PB_OK;
}

pb_status pb_hip_solve_batch(pb_ctx* ctx, const pb_batch_view* batch, const pb_solve_options* options,
                             pb_window_report* reports, pb_solve_report* summary, const volatile int32_t* cancel)
{
    if (!ctx || !summary) return PB_ERR_INVALID;
    std::memset(summary, 0, sizeof(*summary));
    summary->gauge_anchor_camera = summary->gauge_scale_camera = -1;
    auto invalid = [&](pb_status st, const std::string& m) {
        summary->solve_status = st == PB_ERR_UNSUPPORTED ? PB_SOLVE_UNSUPPORTED : PB_SOLVE_INVALID_INPUT;
        setMessage(summary, m);
        return fail(ctx, st, m);
    };
    if (!batch || batch->num_windows <= 0 || !batch->windows) return invalid(PB_ERR_INVALID, "batch has no windows");
    if (!reports) return invalid(PB_ERR_INVALID, "reports is NULL");
    std::memset(reports, 0, sizeof(pb_window_report) * size_t(batch->num_windows));
    const std::string oerr = validateOptions(options);
    if (!oerr.empty()) return invalid(PB_ERR_INVALID, oerr);
    if (!options->use_online_schur)
        return invalid(PB_ERR_UNSUPPORTED, "the MI355X backend implements the online point-Schur path only");
    ctx->profile = false;
    const pb_status st = guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (cancel && *cancel)
        {
            summary->solve_status = PB_SOLVE_CANCELLED;
            setMessage(summary, "cancelled before start");
            return PB_ERR_CANCELLED;
        }
        const auto ts = Clock::now();
        prepareBatch(ctx, batch, *options);
        fillReportCounts(ctx, summary);
        ctx->P.loss = options->robust_loss;
        ctx->P.lossScale = options->robust_loss_scale;
        if (ctx->prep.observations == 0)
        {
            summary->solve_status = PB_SOLVE_INVALID_INPUT;
            setMessage(summary, "no usable tracks");
            return PB_ERR_INVALID;
        }
        const bool mixed = options->precision == PB_PRECISION_MIXED;
        (void)work<double>(ctx);
        if (mixed) (void)work<float>(ctx);
        check(hipStreamSynchronize(ctx->stream), "setup");
        summary->setup_seconds = seconds(ts);
        summary->online_schur_used = 1;
        summary->mixed_precision_used = mixed ? 1 : 0;
        ctx->lastOptions = *options;
        runBatch(ctx, *options, mixed, reports, summary, cancel);
        return summary->solve_status == PB_SOLVE_CANCELLED ? PB_ERR_CANCELLED : PB_OK;
    });
    if (st != PB_OK && st != PB_ERR_CANCELLED && summary->solve_status == PB_SOLVE_NOT_RUN)
    {
        summary->solve_status = st == PB_ERR_UNSUPPORTED  ? PB_SOLVE_UNSUPPORTED
                                : st == PB_ERR_INVALID    ? PB_SOLVE_INVALID_INPUT
                                : st == PB_ERR_DEVICE     ? PB_SOLVE_BACKEND_UNAVAILABLE
                                                          : PB_SOLVE_NUMERICAL_FAILURE;
        setMessage(summary, ctx->error);
    }
    return st;
}

pb_status pb_hip_get_masks(pb_ctx* ctx, int32_t* cameraBlock, int32_t* intrinsicBlock, int32_t* groupOfCam,
                           int32_t* trackBlock, int32_t* activeTrack)
{
    if (!ctx || !ctx->structureValid) return fail(ctx, PB_ERR_INVALID, "no prepared problem (solve first)");
    const Prepared& p = ctx->prep;
    for (int c = 0; c < ctx->in.nCam; ++c)
    {
        if (cameraBlock) cameraBlock[c] = p.rig ? -1 : p.cameraBlock[c];  // a rig's blocks are its captures and sensors
        if (intrinsicBlock) intrinsicBlock[c] = p.intrinsicBlock[c];
        if (groupOfCam) groupOfCam[c] = p.groupOfCam[c];
    }
    for (int t = 0; t < ctx->in.nTracks; ++t)
    {
        if (trackBlock) trackBlock[t] = p.trackBlock[t];
        if (activeTrack) activeTrack[t] = p.activeTrack[t];
    }
    return PB_OK;
}

pb_status pb_hip_kernel_stats(pb_ctx* ctx, const char* name, int64_t* launches, double* total_ms,
                              double* bytes_per_launch)
{
    if (!ctx || !name) return PB_ERR_INVALID;
    for (int f = 0; f < kFamilyCount; ++f)
    {
        if (std::strcmp(kFamilies[f], name) != 0) continue;
        if (launches) *launches = ctx->fam[f].launches;
        if (total_ms) *total_ms = ctx->fam[f].ms;
        if (bytes_per_launch)
        {
            // algorithmic bytes: each array element the kernel family must touch, counted once
            const pbk::Problem& P = ctx->P;
            const double t = ctx->lastScalarBytes;
            double b = 0.0;
            switch (f)
            {
            case 0:  // obs (cam, uv, w, two entry ids) + point X + cross blocks out + L, bp, z out
                b = P.nObs * (4.0 + 16 + 8 + 8) + P.nActive * 4.0 * 2 + P.nElim * (24.0 + 4 + 6 * t + 48) +
                    P.nEntries * 18 * t;
                break;
            case 1:  // camera permutation + obs (point id, uv, w) + point X
                b = P.nObs * (4.0 + 4 + 16 + 8) + P.nActive * 24.0;
                break;
            case 2:  // pairs + cross blocks U + z of the points of diagonal blocks + S out
                b = P.nPairs * 8.0 + P.nEntries * (4.0 + 18 * t) + P.nElim * 24.0 + P.nNnz * 36 * t;
                break;
            case 3:  // S + column ids; per unit the preconditioner block, 17 six-vectors (r, u, w, s, p, x in;
                     // column u, Mw, Ms; r, u, w, Mw, s, Ms, p, x out) and the partials in and out
                b = P.nNnz * (36 * t + 4) + P.nUnits * (36 * t + 17 * 6 * t + 48);
                break;
            case 4:  // obs (cam, point, uv, w) + point X, dp, elim id
                b = P.nObs * (4.0 + 4 + 16 + 8) + P.nActive * (24.0 + 4) + P.nElim * 24.0;
                break;
            case 5:  // cross blocks U + units + point z, L, bp + dp out
                b = P.nEntries * (18 * t + 4) + P.nElim * (24.0 + 6 * t + 24 + 24);
                break;
            case kFamilyDenseFactor:  // floating-point operations of the Cholesky factorisation, n^3 / 3
                b = std::pow(6.0 * P.nUnits, 3) / 3.0;
                break;
            case kFamilyDenseSolve:  // floating-point operations of the two triangular sweeps, 2 n^2
                b = 2.0 * std::pow(6.0 * P.nUnits, 2);
                break;
            case kFamilyBandFactor:  // floating-point operations of the cyclic reduction as planned
                b = ctx->band.view.factorFlops;
                break;
            case kFamilyBandSolve:
                b = ctx->band.view.solveFlops;
                break;
            default:
                break;
            }
            *bytes_per_launch = b;
        }
        return PB_OK;
    }
    return fail(ctx, PB_ERR_INVALID, std::string("unknown kernel family ") + name);
}

pb_status pb_hip_reduced_system(pb_ctx* ctx, const pb_solve_options* options, double damping, double* matrix,
                                double* rhs, double* cost, int32_t* primary_blocks)
{
    if (!ctx || !ctx->uploaded) return PB_ERR_INVALID;
    if (ctx->batchMode) return fail(ctx, PB_ERR_INVALID, "the context holds a batch; upload a problem first");
    const std::string oerr = validateOptions(options);
    if (!oerr.empty()) return fail(ctx, PB_ERR_INVALID, oerr);
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        ensureStructure(ctx, *options);
        applySolveOptions(ctx, *options);
        ctx->profile = false;
        if (options->precision == PB_PRECISION_MIXED)
            return exportReduced<float>(ctx, *options, damping, matrix, rhs, cost, primary_blocks);
        return exportReduced<double>(ctx, *options, damping, matrix, rhs, cost, primary_blocks);
    });
}

pb_status pb_hip_evaluate_cost(pb_ctx* ctx, const pb_solve_options* options, double* cost)
{
    if (!cost) return PB_ERR_INVALID;
    return pb_hip_reduced_system(ctx, options, 0.0, nullptr, nullptr, cost, nullptr);
}

void pb_default_quality_options(pb_quality_options* q)
{
    if (!q) return;
    std::memset(q, 0, sizeof(*q));
    q->enable_point_filter = 1;
    q->filter_max_reproj_error = 2.5;
    q->filter_sigma_factor = 3.0;
    q->enable_quality_gate = 1;
    q->max_accepted_rms_growth = 1.25;
    q->min_accepted_valid_track_ratio = 0.60;
}

pb_status pb_hip_quality(pb_ctx* ctx, const pb_quality_options* q, pb_quality_report* rep, double* rmsBefore,
                         double* rmsAfter, int32_t* validOut)
{
    if (!ctx || !q || !rep) return PB_ERR_INVALID;
    std::memset(rep, 0, sizeof(*rep));
    if (!ctx->uploaded || !ctx->haveResult || ctx->batchMode || !ctx->structureValid)
        return fail(ctx, PB_ERR_INVALID, "pb_hip_quality needs the result of a pb_hip_solve on this context");
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t s = ctx->stream;
        const pbk::Problem& P = ctx->P;
        const Layout& L = ctx->lay;
        const Input& in = ctx->in;
        pbk::Quality& Q = ctx->Q;
        if (!Q.out)
        {
            DeviceAlloc& A = ctx->dquality;
            const size_t n = size_t(std::max(1, P.nActive));
            Q.before = A.alloc<double>(n);
            Q.after = A.alloc<double>(n);
            Q.key = A.alloc<double>(n);
            Q.sorted = A.alloc<double>(n);
            Q.valid = A.alloc<int32_t>(n);
            Q.out = A.alloc<double>(8);
            Q.tempBytes = std::max<size_t>(1, pbk::QualityLauncher::sortScratchBytes(P.nActive));
            Q.temp = A.alloc<char>(Q.tempBytes);
        }
        // effective cameras of the refined state (the solve's last accepted step moved them after the last build)
        pbk::Launcher<double>::prepareCameras(P, s);
        pbk::QualityLauncher::run(P, Q, q->enable_point_filter, q->filter_max_reproj_error, q->filter_sigma_factor, s);
        double out[8] = {};
        check(hipMemcpyAsync(out, Q.out, sizeof(out), hipMemcpyDeviceToHost, s), "readback");
        std::vector<double> before, after;
        std::vector<int32_t> valid;
        const bool perTrack = rmsBefore || rmsAfter || validOut;
        if (perTrack && P.nActive > 0)
        {
            before.resize(P.nActive);
            after.resize(P.nActive);
            valid.resize(P.nActive);
            check(hipMemcpyAsync(before.data(), Q.before, sizeof(double) * P.nActive, hipMemcpyDeviceToHost, s), "readback");
            check(hipMemcpyAsync(after.data(), Q.after, sizeof(double) * P.nActive, hipMemcpyDeviceToHost, s), "readback");
            check(hipMemcpyAsync(valid.data(), Q.valid, sizeof(int32_t) * P.nActive, hipMemcpyDeviceToHost, s), "readback");
        }
        check(hipStreamSynchronize(s), "quality");
        const double inf = std::numeric_limits<double>::infinity();
        const int kept = static_cast<int>(out[3]);
        rep->total_tracks = in.nTracks;
        rep->optimized_tracks = kept;
        rep->mean_rms_before = kept > 0 ? out[4] / kept : inf;
        rep->mean_rms_after = kept > 0 ? out[5] / kept : inf;
        rep->valid_track_ratio = in.nTracks > 0 ? double(kept) / double(in.nTracks) : 0.0;
        rep->median_rms_after = out[1];
        rep->filter_threshold = out[2];
        rep->solve_status = ctx->lastSolveStatus;
        rep->solution_usable = ctx->lastUsable;
        // finalizeBundleAdjustResult (BundleAdjustQuality.cpp:490-501): no kept track downgrades a usable solve
        if (kept == 0 && (rep->solve_status == PB_SOLVE_SUCCESS || rep->solve_status == PB_SOLVE_NO_CONVERGENCE))
        {
            rep->solve_status = PB_SOLVE_NUMERICAL_FAILURE;
            rep->solution_usable = 0;
            std::snprintf(rep->message, sizeof(rep->message), "final quality check kept no valid track");
        }
        // resultFailsQualityGate, image checks (solver.cpp:123-171)
        if (q->enable_quality_gate)
        {
            const char* why = nullptr;
            const double growth = std::max(0.0, q->max_accepted_rms_growth);
            const bool any = in.nTracks > 0;
            if (!rep->solution_usable)
                why = "quality gate rejected a non-usable solution";
            else if (any && !std::isfinite(rep->mean_rms_after))
                why = "quality gate rejected a non-finite final RMS";
            else if (any && rep->valid_track_ratio < std::max(0.0, q->min_accepted_valid_track_ratio))
                why = "quality gate rejected a low valid-track ratio";
            else if (any && growth > 0.0 && std::isfinite(rep->mean_rms_before))
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
        if (perTrack)
        {
            // inactive tracks: input-state RMS on the host (they are never refined), refined RMS inf
            std::vector<pbdev::Camera> inCams;
            if (rmsBefore)
            {
                inCams.resize(in.nCam);
                for (int c = 0; c < in.nCam; ++c) inCams[c] = inputCamera(in, c);
            }
            for (int t = 0; t < in.nTracks; ++t)
            {
                const int a = L.activeIndexOf[t];
                if (a >= 0)
                {
                    if (rmsBefore) rmsBefore[t] = before[a];
                    if (rmsAfter) rmsAfter[t] = after[a];
                    if (validOut) validOut[t] = valid[a];
                    continue;
                }
                if (rmsAfter) rmsAfter[t] = inf;
                if (validOut) validOut[t] = 0;
                if (rmsBefore) rmsBefore[t] = hostStrictRms(in, inCams, t);
            }
        }
        return PB_OK;
    });
}

pb_status pb_hip_set_progress(pb_ctx* ctx, pb_progress_fn fn, void* user)
{
    if (!ctx) return PB_ERR_INVALID;
    ctx->progressFn = fn;
    ctx->progressUser = fn ? user : nullptr;
    return PB_OK;
}

pb_status pb_hip_direct_solve(pb_ctx* ctx, const pb_solve_options* options, double damping, int32_t route,
                              double* x, double* matrix, double* rhs, int32_t* info, double* arena,
                              int64_t arena_capacity)
{
    if (!ctx || !ctx->uploaded || !x) return PB_ERR_INVALID;
    if (ctx->batchMode) return fail(ctx, PB_ERR_INVALID, "the context holds a batch; upload a problem first");
    if (route != PB_REDUCED_DENSE && route != PB_REDUCED_BAND) return fail(ctx, PB_ERR_INVALID, "route must be dense or band");
    if (ctx->in.rig()) return fail(ctx, PB_ERR_UNSUPPORTED, "unsupported: the direct-solve hook takes problems without a rig");
    const std::string oerr = validateOptions(options);
    if (!oerr.empty()) return fail(ctx, PB_ERR_INVALID, oerr);
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        ensureStructure(ctx, *options);
        applySolveOptions(ctx, *options);
        ctx->profile = false;
        hipStream_t s = ctx->stream;
        pbk::Work<double>& W = work<double>(ctx);
        resetState(ctx);
        const BuildResult b = build(ctx, W, damping);
        if (b.invalid) return fail(ctx, PB_ERR_NUMERIC, "invalid projection at the initial state");
        const bool band = route == PB_REDUCED_BAND;
        if (band && !planBand(ctx))
            return fail(ctx, PB_ERR_UNSUPPORTED, "the band route does not fit this reduced system: " + ctx->band.why);
        ctx->band.active = band;
        const pbk::Problem& P = ctx->P;
        const int n = 6 * P.nUnits;
        bool seed = false;
        const PcgResult lin = denseSolve<double>(ctx, W, 0.0, false, &seed, nullptr);
        check(hipStreamSynchronize(s), "direct solve");
        std::vector<double> hx(n, 0.0);
        if (n > 0 && lin.converged)
            check(hipMemcpy(hx.data(), W.x, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
        std::copy(hx.begin(), hx.end(), x);
        if (info)
        {
            std::fill(info, info + 8, 0);
            info[0] = lin.breakdown ? 1 : 0;
            info[1] = int32_t(std::min(2147483647.0, std::max(0.0, lin.rnorm * 1e15)));
            if (band)
            {
                const pbk::BandView& V = ctx->band.view;
                info[2] = V.d.sp;
                info[3] = V.d.m;
                info[4] = V.d.nb;
                info[5] = V.levels;
                info[6] = ctx->band.bandwidth;
            }
            info[7] = n;
        }
        if (matrix || rhs)
        {
            std::vector<double> S(size_t(P.nNnz) * 36), r(n);
            if (!S.empty()) check(hipMemcpy(S.data(), W.S, sizeof(double) * S.size(), hipMemcpyDeviceToHost), "download");
            if (n > 0) check(hipMemcpy(r.data(), W.rhs, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
            if (rhs) std::copy(r.begin(), r.end(), rhs);
            if (matrix)
            {
                std::fill(matrix, matrix + size_t(n) * n, 0.0);
                for (int u = 0; u < P.nUnits; ++u)
                    for (int k = ctx->lay.rows.rowPtr[u]; k < ctx->lay.rows.rowPtr[u + 1]; ++k)
                    {
                        const int v = ctx->lay.rows.colIdx[k];
                        for (int e = 0; e < 36; ++e)
                            matrix[size_t(6 * v + e % 6) * n + 6 * u + e / 6] = S[size_t(k) * 36 + e];
                    }
            }
        }
        if (arena && band && ctx->band.Z64)
        {
            const long long size = ctx->band.view.d.size;
            if (arena_capacity < size) return fail(ctx, PB_ERR_INVALID, "arena capacity below the band arena size");
            check(hipMemcpy(arena, ctx->band.Z64, sizeof(double) * size_t(size), hipMemcpyDeviceToHost), "download");
        }
        return PB_OK;
    });
}

pb_status pb_hip_dense_solve(pb_ctx* ctx, int32_t n, int32_t fp32, const double* A, const double* b, double* x,
                             int32_t* info, double* seconds)
{
    if (!ctx || n <= 0 || !A || !b || !x || (fp32 != 0 && fp32 != 1)) return fail(ctx, PB_ERR_INVALID, "invalid dense solve");
    return guarded(ctx, [&]() -> pb_status {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t s = ctx->stream;
        const int nt = (n + pbk::kDenseTile - 1) / pbk::kDenseTile;
        const size_t npad = size_t(nt) * pbk::kDenseTile;
        ensureDense(ctx, nt, fp32 != 0);
        auto& D = ctx->dense;
        // the solver works on the leading npad x npad block of the matrix buffer (leading dimension npad)
        struct Events
        {
            hipEvent_t e[3] = {nullptr, nullptr, nullptr};
            ~Events()
            {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } evs;
        hipEvent_t* ev = evs.e;
        for (int i = 0; i < 3; ++i) check(hipEventCreate(&ev[i]), "hipEventCreate");
        auto run = [&](auto* Adev, auto* vdev) {
            using TA = std::remove_pointer_t<decltype(Adev)>;
            std::vector<TA> hA(npad * npad, TA(0)), hb(npad, TA(0));
            for (size_t c = 0; c < npad; ++c)
                for (size_t r = 0; r < npad; ++r)
                    hA[c * npad + r] = (r < size_t(n) && c < size_t(n)) ? TA(A[c * n + r]) : TA(r == c ? 1 : 0);
            for (int i = 0; i < n; ++i) hb[i] = TA(b[i]);
            check(hipMemcpyAsync(Adev, hA.data(), sizeof(TA) * hA.size(), hipMemcpyHostToDevice, s), "upload");
            check(hipMemcpyAsync(vdev, hb.data(), sizeof(TA) * hb.size(), hipMemcpyHostToDevice, s), "upload");
            check(hipEventRecord(ev[0], s), "hipEventRecord");
            pbk::DenseLauncher<TA>::factor(Adev, nt, D.st, s);
            check(hipEventRecord(ev[1], s), "hipEventRecord");
            pbk::DenseLauncher<TA>::solve(Adev, nt, vdev, D.st, s);
            check(hipEventRecord(ev[2], s), "hipEventRecord");
            check(hipMemcpyAsync(hb.data(), vdev, sizeof(TA) * npad, hipMemcpyDeviceToHost, s), "download");
            int32_t words[pbk::kDenseWordCount];
            check(hipMemcpyAsync(words, D.st.words, sizeof(words), hipMemcpyDeviceToHost, s), "download");
            check(hipStreamSynchronize(s), "dense solve");
            if (words[pbk::kDenseTimeout]) throw std::runtime_error("dense triangular solve stopped waiting for a tile");
            for (int i = 0; i < n; ++i) x[i] = double(hb[i]);
            if (info) *info = words[pbk::kDenseNotSpd] ? 1 : 0;
        };
        if (fp32)
            run(D.A32, D.v32);
        else
            run(D.A64, D.v64);
        float f = 0.0f, t = 0.0f;
        check(hipEventElapsedTime(&f, ev[0], ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&t, ev[1], ev[2]), "hipEventElapsedTime");
        if (seconds)
        {
            seconds[0] = f * 1e-3;
            seconds[1] = t * 1e-3;
        }
        return PB_OK;
    });
}

const char* pb_hip_last_error(const pb_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

}  // extern "C"

// Context access for pb_hip_dense_solve_noalloc (pb_noalloc.hip): its device, its stream and the pinned readback
// area allocated when the context was created.
bool pbNoallocContext(pb_ctx* ctx, int* device, hipStream_t* stream, int32_t** pinnedWords)
{
    if (!ctx || !ctx->stream || !ctx->pinned) return false;
    *device = ctx->device;
    *stream = ctx->stream;
    *pinnedWords = reinterpret_cast<int32_t*>(ctx->pinned + kPinDense);
    return true;
}

pb_status pbNoallocFail(pb_ctx* ctx, pb_status st, const std::string& message) { return fail(ctx, st, message); }
