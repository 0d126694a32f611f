// pb_outcome.cpp - id-keyed outcome validation and comparison (host code, no device work). Restates
// validatePhotoAlignmentOutcome and comparePhotoAlignmentOutcomes (src/photo_alignment.cpp:193-318) over the
// C ABI's borrowed arrays.
#include "../../include/pb_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

namespace
{
    constexpr double kDegreesPerRadian = 180.0 / 3.14159265358979323846;

    bool isFinite(double v) { return std::isfinite(v); }

    void copyMessage(char* dst, int32_t size, const std::string& msg)
    {
        if (dst && size > 0) std::snprintf(dst, static_cast<size_t>(size), "%s", msg.c_str());
    }

    bool makeIndex(const char* const* ids, int32_t n, const char* label, std::map<std::string, int>* index,
                   std::string* error)
    {
        index->clear();
        if (n > 0 && !ids)
        {
            *error = std::string(label) + " ids are NULL";
            return false;
        }
        for (int32_t i = 0; i < n; ++i)
        {
            const std::string id = ids[i] ? ids[i] : "";
            if (id.empty())
            {
                *error = std::string(label) + " id must not be empty";
                return false;
            }
            if (!index->emplace(id, i).second)
            {
                *error = std::string(label) + " ids must be unique: " + id;
                return false;
            }
        }
        return true;
    }

    bool validate(const pb_outcome_view* o, std::string* error)
    {
        if (!o)
        {
            *error = "outcome is NULL";
            return false;
        }
        if (!isFinite(o->mean_rms_before) || !isFinite(o->mean_rms_after))
        {
            *error = "photo-alignment outcome RMS values must be finite";
            return false;
        }
        if (o->num_cameras < 0 || o->num_tracks < 0)
        {
            *error = "negative counts";
            return false;
        }
        if ((o->num_cameras > 0 && (!o->cam_rotation || !o->cam_center || !o->cam_intrinsics || !o->cam_image_size)) ||
            (o->num_tracks > 0 && (!o->track_point || !o->track_valid || !o->track_rms_before || !o->track_rms_after)))
        {
            *error = "outcome arrays are NULL";
            return false;
        }
        std::map<std::string, int> cams, tracks;
        if (!makeIndex(o->camera_ids, o->num_cameras, "camera", &cams, error) ||
            !makeIndex(o->track_ids, o->num_tracks, "track", &tracks, error))
            return false;
        for (const auto& [id, c] : cams)
        {
            const double* s = o->cam_intrinsics + 13 * c;
            const double fx = s[0], fy = s[0] * std::exp(s[1]);
            const double* C = o->cam_center + 3 * c;
            const bool valid = o->cam_image_size[2 * c] > 0 && o->cam_image_size[2 * c + 1] > 0 && isFinite(fx) &&
                               isFinite(fy) && fx != 0.0 && fy != 0.0 && isFinite(C[0]) && isFinite(C[1]) && isFinite(C[2]);
            if (!valid)
            {
                *error = "photo-alignment camera " + id + " is invalid";
                return false;
            }
        }
        for (const auto& [id, t] : tracks)
        {
            const double* X = o->track_point + 3 * t;
            if (!isFinite(X[0]) || !isFinite(X[1]) || !isFinite(X[2]) || !isFinite(o->track_rms_before[t]) ||
                !isFinite(o->track_rms_after[t]))
            {
                *error = "photo-alignment track " + id + " contains non-finite state";
                return false;
            }
        }
        return true;
    }

    double rotationDifferenceDegrees(const double* a, const double* b)
    {
        double trace = 0.0;  // trace(a^T b)
        for (int i = 0; i < 9; ++i) trace += a[i] * b[i];
        return std::acos(std::clamp(0.5 * (trace - 1.0), -1.0, 1.0)) * kDegreesPerRadian;
    }
}  // namespace

extern "C" {

void pb_default_comparison_tolerance(pb_comparison_tolerance* t)
{
    if (!t) return;
    t->rms = 1e-8;
    t->camera_center_meters = 1e-8;
    t->camera_rotation_degrees = 1e-8;
    t->point_meters = 1e-8;
    t->require_termination_status_match = 1;
    t->require_termination_reason_match = 1;
    t->require_requested_backend_match = 1;
    t->require_used_backend_match = 1;
}

pb_status pb_validate_outcome(const pb_outcome_view* outcome, char* error, int32_t error_size)
{
    std::string err;
    const bool ok = validate(outcome, &err);
    copyMessage(error, error_size, ok ? std::string() : err);
    return ok ? PB_OK : PB_ERR_INVALID;
}

pb_status pb_compare_outcomes(const pb_outcome_view* ref, const pb_outcome_view* cand, const pb_comparison_tolerance* tol,
                              pb_comparison* cmp, double* centerDiff, double* rotationDiff, double* pointDiff,
                              int32_t* validMismatch)
{
    if (!cmp || !tol) return PB_ERR_INVALID;
    std::memset(cmp, 0, sizeof(*cmp));
    if (!isFinite(tol->rms) || tol->rms < 0.0 || !isFinite(tol->camera_center_meters) || tol->camera_center_meters < 0.0 ||
        !isFinite(tol->camera_rotation_degrees) || tol->camera_rotation_degrees < 0.0 || !isFinite(tol->point_meters) ||
        tol->point_meters < 0.0)
    {
        copyMessage(cmp->message, sizeof(cmp->message), "photo-alignment comparison tolerances must be finite and non-negative");
        return PB_ERR_INVALID;
    }
    std::string err;
    if (!validate(ref, &err) || !validate(cand, &err))
    {
        copyMessage(cmp->message, sizeof(cmp->message), err);
        return PB_ERR_INVALID;
    }
    std::map<std::string, int> rc, cc, rt, ct;
    makeIndex(ref->camera_ids, ref->num_cameras, "camera", &rc, &err);
    makeIndex(cand->camera_ids, cand->num_cameras, "camera", &cc, &err);
    makeIndex(ref->track_ids, ref->num_tracks, "track", &rt, &err);
    makeIndex(cand->track_ids, cand->num_tracks, "track", &ct, &err);
    auto sameIds = [](const std::map<std::string, int>& a, const std::map<std::string, int>& b) {
        return a.size() == b.size() &&
               std::equal(a.begin(), a.end(), b.begin(), [](const auto& x, const auto& y) { return x.first == y.first; });
    };
    if (!sameIds(rc, cc) || !sameIds(rt, ct))
    {
        copyMessage(cmp->message, sizeof(cmp->message), "camera or track id sets differ");
        return PB_ERR_INVALID;
    }
    cmp->comparable = 1;
    const std::string rr = ref->termination_reason ? ref->termination_reason : "";
    const std::string cr = cand->termination_reason ? cand->termination_reason : "";
    cmp->termination_status_matches = ref->termination_status == cand->termination_status;
    cmp->termination_reason_matches = rr == cr;
    cmp->requested_backend_matches = ref->requested_backend == cand->requested_backend;
    cmp->used_backend_matches = ref->used_backend == cand->used_backend;
    cmp->rms_before_difference = std::abs(ref->mean_rms_before - cand->mean_rms_before);
    cmp->rms_after_difference = std::abs(ref->mean_rms_after - cand->mean_rms_after);
    for (const auto& [id, i] : rc)
    {
        const int j = cc.at(id);
        const double* a = ref->cam_center + 3 * i;
        const double* b = cand->cam_center + 3 * j;
        const double d = std::hypot(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
        const double r = rotationDifferenceDegrees(ref->cam_rotation + 9 * i, cand->cam_rotation + 9 * j);
        cmp->maximum_camera_center_difference_meters = std::max(cmp->maximum_camera_center_difference_meters, d);
        cmp->maximum_camera_rotation_difference_degrees = std::max(cmp->maximum_camera_rotation_difference_degrees, r);
        if (centerDiff) centerDiff[i] = d;
        if (rotationDiff) rotationDiff[i] = r;
    }
    for (const auto& [id, i] : rt)
    {
        const int j = ct.at(id);
        const bool va = ref->track_valid[i] != 0, vb = cand->track_valid[j] != 0;
        double d = 0.0;
        if (va != vb) ++cmp->valid_track_mask_mismatches;
        if (va && vb)
        {
            const double* a = ref->track_point + 3 * i;
            const double* b = cand->track_point + 3 * j;
            d = std::hypot(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
            cmp->maximum_point_difference_meters = std::max(cmp->maximum_point_difference_meters, d);
        }
        if (pointDiff) pointDiff[i] = d;
        if (validMismatch) validMismatch[i] = va != vb ? 1 : 0;
    }
    const bool statusOk = !tol->require_termination_status_match || cmp->termination_status_matches;
    const bool reasonOk = !tol->require_termination_reason_match || cmp->termination_reason_matches;
    const bool requestedOk = !tol->require_requested_backend_match || cmp->requested_backend_matches;
    const bool usedOk = !tol->require_used_backend_match || cmp->used_backend_matches;
    cmp->equivalent = statusOk && reasonOk && requestedOk && usedOk && cmp->rms_before_difference <= tol->rms &&
                      cmp->rms_after_difference <= tol->rms &&
                      cmp->maximum_camera_center_difference_meters <= tol->camera_center_meters &&
                      cmp->maximum_camera_rotation_difference_degrees <= tol->camera_rotation_degrees &&
                      cmp->maximum_point_difference_meters <= tol->point_meters &&
                      cmp->valid_track_mask_mismatches == 0;
    copyMessage(cmp->message, sizeof(cmp->message),
                cmp->equivalent ? "photo-alignment outcomes are equivalent" : "photo-alignment outcomes differ");
    return PB_OK;
}

}  // extern "C"
