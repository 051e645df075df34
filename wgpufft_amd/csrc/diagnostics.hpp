// Structured failure diagnostics (blockers), shared by the C ABI and the
// plan diagnostics document.
#pragma once
#include <cstdint>
#include <string>

#include "../../include/mifft.h"

namespace mifft {

// FftBlocker (src/diagnostics.rs:36-59).
struct Blocker {
    std::string kind, route, stage, layout, helper, reason;
    uint64_t required = 0, actual = 0, limit = 0;
    bool has_required = false, has_actual = false, has_limit = false;
    Blocker& set(const std::string& k, const std::string& r, const std::string& s, const std::string& l) {
        kind = k;
        route = r;
        stage = s;
        layout = l;
        return *this;
    }
    Blocker& req(uint64_t v) {
        required = v;
        has_required = true;
        return *this;
    }
    Blocker& act(uint64_t v) {
        actual = v;
        has_actual = true;
        return *this;
    }
    Blocker& lim(uint64_t v) {
        limit = v;
        has_limit = true;
        return *this;
    }
};

std::string blocker_json(const Blocker& b);
// The FftError variant name of a status ("InvalidAxis", ...).
const char* status_variant(mifft_status s);
std::string tuning_json(const mifft_tuning& t);
Blocker creation_blocker(const mifft_config* c, mifft_kind kind, mifft_status code, const std::string& reason);
// FftPlanCreationError diagnostics of a failed plan creation.
std::string creation_failure_json(const mifft_config* c, mifft_kind kind, mifft_status code, const std::string& detail);

}  // namespace mifft
