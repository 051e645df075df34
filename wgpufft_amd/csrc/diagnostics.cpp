// Structured failure diagnostics: FftError::diagnostics (src/error.rs:498-850)
// and the blockers of FftPlan::diagnostics_for_workspace / _for_views
// (src/plan.rs:1023-1109, 1396-1500), as JSON.  A blocker names what stops a
// plan or an execution: its kind (the reference's FftBlockerKind), route,
// stage, layout, helper buffer, reason and byte figures.
#include <sstream>

#include "diagnostics.hpp"

namespace mifft {

static std::string jstr(const std::string& v) {
    std::string o = "\"";
    for (char c : v) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) continue;
        o += c;
    }
    return o + "\"";
}

static std::string jopt(const std::string& v) { return v.empty() ? "null" : jstr(v); }

std::string blocker_json(const Blocker& b) {
    std::ostringstream d;
    d << "{\"kind\":" << jstr(b.kind) << ",\"route\":" << jopt(b.route) << ",\"stage\":" << jopt(b.stage)
      << ",\"layout\":" << jopt(b.layout) << ",\"helper_buffer\":" << jopt(b.helper) << ",\"reason\":"
      << jstr(b.reason) << ",\"required_bytes\":";
    if (b.has_required) d << b.required; else d << "null";
    d << ",\"actual_bytes\":";
    if (b.has_actual) d << b.actual; else d << "null";
    d << ",\"limit_bytes\":";
    if (b.has_limit) d << b.limit; else d << "null";
    d << "}";
    return d.str();
}

const char* status_variant(mifft_status s) {
    switch (s) {
        case MIFFT_SUCCESS: return "Success";
        case MIFFT_ERR_ZERO_LENGTH: return "ZeroLength";
        case MIFFT_ERR_ZERO_BATCH: return "ZeroBatch";
        case MIFFT_ERR_LENGTH_TOO_LARGE: return "LengthTooLarge";
        case MIFFT_ERR_UNSUPPORTED_LENGTH: return "UnsupportedLength";
        case MIFFT_ERR_PRECISION_UNSUPPORTED: return "PrecisionUnsupported";
        case MIFFT_ERR_INVALID_TUNING: return "InvalidTuning";
        case MIFFT_ERR_EMPTY_AXES: return "EmptyAxes";
        case MIFFT_ERR_INVALID_AXIS: return "InvalidAxis";
        case MIFFT_ERR_DUPLICATE_AXIS: return "DuplicateAxis";
        case MIFFT_ERR_INVALID_REAL_DIRECTION: return "InvalidRealTransformDirection";
        case MIFFT_ERR_UNSUPPORTED_REAL_AXES: return "UnsupportedRealAxes";
        case MIFFT_ERR_BUFFER_TOO_SMALL: return "BufferViewTooSmall";
        case MIFFT_ERR_WORKSPACE_TOO_SMALL: return "WorkspaceTooSmall";
        case MIFFT_ERR_IN_PLACE_UNSUPPORTED: return "InPlaceUnsupported";
        case MIFFT_ERR_ALLOCATION_FAILED: return "GpuPlanResourceAllocationFailed";
        case MIFFT_ERR_HOST_BUFFER_LENGTH: return "HostBufferLength";
        case MIFFT_ERR_NULL_ARGUMENT: return "NullArgument";
        case MIFFT_ERR_INVALID_ENUM: return "InvalidEnum";
        case MIFFT_ERR_RANK_TOO_LARGE: return "RankTooLarge";
        case MIFFT_ERR_HIP: return "Hip";
        case MIFFT_ERR_RECORDER_STATE: return "RecorderState";
        case MIFFT_ERR_OVER_BUDGET: return "OverBudget";
        case MIFFT_ERR_BUSY: return "Busy";
        case MIFFT_ERR_UNALIGNED: return "BufferViewOffsetUnaligned";
        case MIFFT_ERR_LAYOUT_ZERO_STRIDE: return "BufferLayoutZeroStride";
        case MIFFT_ERR_LAYOUT_BATCH_STRIDE: return "BufferLayoutBatchStrideTooSmall";
        case MIFFT_ERR_LAYOUT_OUT_OF_BOUNDS: return "BufferLayoutOutOfBounds";
        case MIFFT_ERR_LAYOUT_TOO_LARGE: return "BufferLayoutTooLarge";
    }
    return "Unknown";
}

static const char* kind_name(mifft_kind k) { return k == MIFFT_R2C ? "r2c" : k == MIFFT_C2R ? "c2r" : "c2c"; }

// The blocker FftError::diagnostics attaches to a plan-creation error
// (error.rs:521-850), with the figures taken from the configuration.
Blocker creation_blocker(const mifft_config* c, mifft_kind kind, mifft_status code, const std::string& reason) {
    Blocker b;
    b.reason = reason;
    const int rank = c ? c->rank : 0;
    switch (code) {
        case MIFFT_ERR_ZERO_LENGTH:
            b.set("Validation", "plan-config", "config-shape", "length").req(1).act(0);
            break;
        case MIFFT_ERR_ZERO_BATCH:
            b.set("Validation", "plan-config", "config-batch", "batch").req(1).act(0);
            break;
        case MIFFT_ERR_LENGTH_TOO_LARGE: {
            unsigned __int128 t = c ? c->batch : 0;
            for (int i = 0; c && i < rank && i < MIFFT_MAX_RANK; i++) {
                t *= c->shape[i];
                if (t > (unsigned __int128)~0ull) t = (unsigned __int128)~0ull;
            }
            b.set("DeviceLimit", "plan-config", "dispatch-dimensions", "workgroup grid").req((uint64_t)t).lim(0xFFFFFFFFull);
            break;
        }
        case MIFFT_ERR_EMPTY_AXES:
            b.set("Validation", "axis-policy", "axis-policy", "axes").req(1).act(0);
            break;
        case MIFFT_ERR_INVALID_AXIS: {
            int bad = 0;
            for (int k = 0; c && k < c->naxes && k < MIFFT_MAX_RANK; k++)
                if (c->axes[k] < 0 || c->axes[k] >= rank) {
                    bad = c->axes[k];
                    break;
                }
            b.set("Validation", "axis-policy", "axis-policy", "axes").req((uint64_t)rank).act((uint64_t)(int64_t)bad);
            break;
        }
        case MIFFT_ERR_DUPLICATE_AXIS: {
            int dup = 0;
            for (int k = 0; c && k < c->naxes && k < MIFFT_MAX_RANK; k++)
                for (int j = 0; j < k; j++)
                    if (c->axes[j] == c->axes[k]) dup = c->axes[k];
            b.set("Validation", "axis-policy", "axis-" + std::to_string(dup), "axes");
            break;
        }
        case MIFFT_ERR_UNSUPPORTED_LENGTH: {
            // a selected length-1 axis of a transform of more than one point
            // (config.rs:203-205), else the first selected axis
            uint64_t len = c && c->naxes > 0 && c->axes[0] >= 0 && c->axes[0] < rank ? c->shape[c->axes[0]] : 0;
            for (int k = 0; c && k < c->naxes && k < MIFFT_MAX_RANK; k++)
                if (c->axes[k] >= 0 && c->axes[k] < rank && c->shape[c->axes[k]] == 1) len = 1;
            b.set("Unsupported", "mixed-radix", "mixed-radix-factorization", "length").req(len);
            break;
        }
        case MIFFT_ERR_PRECISION_UNSUPPORTED: {
            static const char* pn[] = {"f32", "f64", "df64"};
            const int pr = c && c->precision >= 0 && c->precision <= 2 ? c->precision : 0;
            b.set("Unsupported", kind_name(kind), "precision-capability", std::string("complex-") + pn[pr]);
            break;
        }
        case MIFFT_ERR_INVALID_TUNING: {
            // the detail names the offending field first ("force_rader_axes: ...")
            std::string field = "tuning";
            const size_t colon = reason.find(':');
            if (colon != std::string::npos && reason.find_first_not_of("abcdefghijklmnopqrstuvwxyz_") == colon)
                field = reason.substr(0, colon);
            b.set("Validation", "plan-tuning", "tuning-validation", field);
            break;
        }
        case MIFFT_ERR_INVALID_REAL_DIRECTION:
            b.set("Unsupported", kind_name(kind), "direction", "transform-direction");
            break;
        case MIFFT_ERR_UNSUPPORTED_REAL_AXES:
            b.set("Unsupported", "real", "real-axis-policy", "real-axes").req((uint64_t)rank).act((uint64_t)(c ? c->naxes : 0));
            break;
        case MIFFT_ERR_ALLOCATION_FAILED:
        case MIFFT_ERR_HIP:
            b.set("DeviceLimit", kind_name(kind), "plan-resource-allocation", "GPU resource allocation");
            b.helper = "plan-tables";
            break;
        case MIFFT_ERR_RANK_TOO_LARGE:
            b.set("Validation", "plan-config", "config-shape", "rank").act((uint64_t)(int64_t)rank).lim(MIFFT_MAX_RANK);
            break;
        default:
            b.set("Validation", "plan-config", "config", "arguments");
            break;
    }
    return b;
}

static const char* route_transform(const std::string& route, mifft_kind kind) {
    if (route == "real") return "real";
    if (route == "logical-io") return "unknown";
    return kind_name(kind);
}

std::string tuning_json(const mifft_tuning& t) {
    std::ostringstream d;
    d << "{\"workgroup_size\":" << t.workgroup_size << ",\"fused_workgroup_size\":" << t.fused_workgroup_size
      << ",\"rader_max_prime\":" << t.rader_max_prime << ",\"direct_max_prime\":" << t.direct_max_prime
      << ",\"fuse_long_axes\":" << (t.fuse_long_axes ? "true" : "false")
      << ",\"fuse_small_volumes\":" << (t.fuse_small_volumes ? "true" : "false")
      << ",\"fused_min_convolution_length\":" << t.fused_min_convolution_length << ",\"force_rader_axes\":[";
    for (int i = 0; i < t.n_force_rader_axes && i < MIFFT_MAX_FORCED_AXES; i++) d << (i ? "," : "") << t.force_rader_axes[i];
    d << "],\"force_bluestein_axes\":[";
    for (int i = 0; i < t.n_force_bluestein_axes && i < MIFFT_MAX_FORCED_AXES; i++)
        d << (i ? "," : "") << t.force_bluestein_axes[i];
    d << "]}";
    return d.str();
}

// FftPlanCreationError::from_error_with_tuning (error.rs:220-228): the
// error's diagnostics with the requested tuning.
std::string creation_failure_json(const mifft_config* c, mifft_kind kind, mifft_status code, const std::string& detail) {
    const Blocker b = creation_blocker(c, kind, code, detail);
    std::ostringstream d;
    d << "{\"error\":" << jstr(status_variant(code)) << ",\"route\":{\"transform\":"
      << jstr(route_transform(b.route, kind)) << ",\"route\":" << jstr(b.route)
      << ",\"large_route_mode\":null,\"execution_kind\":null,\"reason_codes\":[],\"attempted_routes\":[]},"
      << "\"stages\":[],\"blockers\":[" << blocker_json(b) << "],\"buffer_requirements\":[],\"active_tuning\":";
    if (c) {
        d << tuning_json(c->tuning);
    } else {
        d << "null";
    }
    d << ",\"device_limits\":null}";
    return d.str();
}

}  // namespace mifft
