// Planner and executor of the host runtime.
//
// Plan creation restates the reference's route choice (FftPlan::c2c ->
// C2cPlan::new -> select_route / AxisPlan::new, src/runtime/c2c.rs:869-1275,
// 7631-7650, src/runtime/axis_plan.rs:516-1144) for gfx950: each transformed
// axis becomes one or more kernel launches ("ops") that read and write device
// buffers named by role (input, output, workspace).  Tables (twiddles,
// chirps, Rader/Bluestein filter spectra) are generated on the device in FP64
// at plan creation and rounded once.
#include "plan.hpp"
#include "diagnostics.hpp"

#include <algorithm>

#include "batch_kernels.hpp"
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>

namespace mifft {

void fail(mifft_status code, const std::string& detail) { throw Error{code, detail}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(MIFFT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

Plan::~Plan() {
    for (DevBuf& b : owned)
        if (b.p) (void)hipFree(b.p);
    if (own_ws.p) (void)hipFree(own_ws.p);
    if (in_copy.p) (void)hipFree(in_copy.p);
    if (stage.p) (void)hipFree(stage.p);
}

// ------------------------------------------------------------ integer helpers
static const uint64_t kFactorOrder[8] = {13, 11, 8, 7, 5, 4, 3, 2};

// factor_supported_length (src/runtime/mod.rs:33-54): greedy over [13,11,8,7,5,4,3,2].
std::vector<uint64_t> factor_supported(uint64_t n, bool* ok) {
    std::vector<uint64_t> f;
    *ok = n > 0;
    uint64_t r = n;
    while (*ok && r > 1) {
        bool found = false;
        for (uint64_t c : kFactorOrder) {
            if (r % c == 0) {
                f.push_back(c);
                r /= c;
                found = true;
                break;
            }
        }
        if (!found) *ok = false;
    }
    return f;
}

static bool is_smooth(uint64_t n) {
    if (n == 0) return false;
    for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull})
        while (n % p == 0) n /= p;
    return n == 1;
}

bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    if (n % 2 == 0) return n == 2;
    for (uint64_t d = 3; d <= n / d; d += 2)
        if (n % d == 0) return false;
    return true;
}

uint64_t mod_pow_u64(uint64_t b, uint64_t e, uint64_t m) {
    unsigned __int128 r = 1 % m, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

// primitive_root_prime (src/runtime/axis_policy.rs:134-153).
uint64_t primitive_root(uint64_t p) {
    if (!is_prime_u64(p)) return 0;
    if (p == 2) return 1;
    std::vector<uint64_t> fs;
    uint64_t n = p - 1;
    for (uint64_t d = 2; d <= n / d; d += (d == 2 ? 1 : 2)) {
        if (n % d == 0) {
            fs.push_back(d);
            while (n % d == 0) n /= d;
        }
    }
    if (n > 1) fs.push_back(n);
    for (uint64_t g = 2; g < p; g++) {
        bool ok = true;
        for (uint64_t f : fs)
            if (mod_pow_u64(g, (p - 1) / f, p) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
    return 0;
}

// next_smooth_at_least (src/runtime/axis_policy.rs:155-161).
uint64_t next_smooth(uint64_t n) {
    uint64_t c = n ? n : 1;
    while (!is_smooth(c)) c++;
    return c;
}

// rader_convolution_length (src/runtime/rader_axis.rs:1270-1289).
uint64_t rader_conv_len(uint64_t p) {
    const uint64_t l = p - 1;
    if (l >= 2 && is_smooth(l)) return l;
    return next_smooth(2 * l - 1);
}

// bluestein_convolution_length (src/runtime/bluestein_axis.rs:1035-1041).
uint64_t bluestein_conv_len(uint64_t n) { return next_smooth(2 * n - 1); }

// ------------------------------------------------------ config validation
static void validate_tuning(const mifft_config& c) {
    const mifft_tuning& t = c.tuning;
    auto pow2 = [](uint32_t v) { return v != 0 && (v & (v - 1)) == 0; };
    if (!pow2(t.workgroup_size))
        fail(MIFFT_ERR_INVALID_TUNING, "workgroup_size must be a non-zero power of two");
    if (!pow2(t.fused_workgroup_size))
        fail(MIFFT_ERR_INVALID_TUNING, "fused_workgroup_size must be a non-zero power of two");
    if (t.workgroup_size > 1024 || t.fused_workgroup_size > 1024)
        fail(MIFFT_ERR_INVALID_TUNING, "workgroup size exceeds the 1024-thread device limit");
    if (t.rader_max_prime < 2) fail(MIFFT_ERR_INVALID_TUNING, "rader_max_prime must be at least 2");
    if (t.n_force_rader_axes < 0 || t.n_force_rader_axes > MIFFT_MAX_FORCED_AXES ||
        t.n_force_bluestein_axes < 0 || t.n_force_bluestein_axes > MIFFT_MAX_FORCED_AXES)
        fail(MIFFT_ERR_INVALID_TUNING, "forced axis count out of range");
    for (int i = 0; i < t.n_force_rader_axes; i++)
        for (int j = 0; j < t.n_force_bluestein_axes; j++)
            if (t.force_rader_axes[i] == t.force_bluestein_axes[j])
                fail(MIFFT_ERR_INVALID_TUNING,
                     "the same axis cannot be forced to both Rader and Bluestein");
    for (int pass = 0; pass < 2; pass++) {
        const int32_t* ax = pass == 0 ? t.force_rader_axes : t.force_bluestein_axes;
        const int n = pass == 0 ? t.n_force_rader_axes : t.n_force_bluestein_axes;
        const char* field = pass == 0 ? "force_rader_axes" : "force_bluestein_axes";
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < i; j++)
                if (ax[j] == ax[i]) fail(MIFFT_ERR_INVALID_TUNING, std::string(field) + ": axis appears more than once");
            if (ax[i] < 0 || ax[i] >= c.rank)
                fail(MIFFT_ERR_INVALID_TUNING, std::string(field) + ": axis is outside the configured shape rank");
            bool selected = false;
            for (int k = 0; k < c.naxes; k++) selected |= c.axes[k] == ax[i];
            if (!selected)
                fail(MIFFT_ERR_INVALID_TUNING, std::string(field) + ": axis is not present in FftConfig::axes");
            const uint64_t len = c.shape[ax[i]];
            if (pass == 0 && (len < 3 || !is_prime_u64(len)))
                fail(MIFFT_ERR_INVALID_TUNING, "forced Rader axes must have prime length at least 3");
            if (pass == 1 && len < 2)
                fail(MIFFT_ERR_INVALID_TUNING, "forced Bluestein axes must have length at least 2");
        }
    }
}

// prod(shape) * batch.  One transform is limited to u32::MAX elements (the
// reference's dispatch limit, config.rs:218-229); more batches than that span
// run as a large-chunk plan of independent batch chunks, up to 2^59 elements
// (byte offsets of 16-byte elements stay within 63 bits).
uint64_t config_total(const mifft_config& c) {
    unsigned __int128 per = 1;
    for (int i = 0; i < c.rank; i++) {
        per *= c.shape[i];
        if (per > 0xFFFFFFFFull)
            fail(MIFFT_ERR_LENGTH_TOO_LARGE, "prod(shape) exceeds u32::MAX elements in one transform");
    }
    const unsigned __int128 t = per * c.batch;
    if (t > 0xFFFFFFFFull && c.batch < 2)
        fail(MIFFT_ERR_LENGTH_TOO_LARGE, "prod(shape)*batch exceeds u32::MAX elements");
    if (t > (1ull << 59)) fail(MIFFT_ERR_LENGTH_TOO_LARGE, "prod(shape)*batch exceeds 2^59 elements");
    return (uint64_t)t;
}

// FftConfig::validate (src/config.rs:178-211).
void validate_config(const mifft_config& c) {
    if (c.rank < 0 || c.rank > MIFFT_MAX_RANK) fail(MIFFT_ERR_RANK_TOO_LARGE, "rank out of range 1..8");
    if (c.rank == 0) fail(MIFFT_ERR_ZERO_LENGTH, "empty shape");
    for (int i = 0; i < c.rank; i++)
        if (c.shape[i] == 0) fail(MIFFT_ERR_ZERO_LENGTH, "zero-length axis");
    if (c.batch == 0) fail(MIFFT_ERR_ZERO_BATCH, "batch must be at least 1");
    if (c.naxes <= 0) fail(MIFFT_ERR_EMPTY_AXES, "no axes selected");
    if (c.naxes > MIFFT_MAX_RANK) fail(MIFFT_ERR_RANK_TOO_LARGE, "too many axes");
    if (c.direction < 0 || c.direction > 1 || c.normalization < 0 || c.normalization > 3 ||
        c.precision < 0 || c.precision > 2)
        fail(MIFFT_ERR_INVALID_ENUM, "direction/normalization/precision out of range");
    bool seen[MIFFT_MAX_RANK] = {false};
    for (int k = 0; k < c.naxes; k++) {
        const int a = c.axes[k];
        if (a < 0 || a >= c.rank)
            fail(MIFFT_ERR_INVALID_AXIS, "axis " + std::to_string(a) + " >= rank " + std::to_string(c.rank));
        if (seen[a]) fail(MIFFT_ERR_DUPLICATE_AXIS, "axis " + std::to_string(a) + " repeated");
        seen[a] = true;
        if (c.shape[a] == 1 && config_total(c) != 1)
            fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "length-1 axis in a transform of more than one point");
    }
    config_total(c);
    validate_tuning(c);
}

// FftConfig::scale_f64 (src/config.rs:274-286): 1/prod(shape), not of the selected axes.
double config_scale(const mifft_config& c) {
    double len = 1.0;
    for (int i = 0; i < c.rank; i++) len *= (double)c.shape[i];
    const bool inv = c.direction == MIFFT_INVERSE;
    switch (c.normalization) {
        case MIFFT_NORM_NONE: return 1.0;
        case MIFFT_NORM_FORWARD: return inv ? 1.0 : 1.0 / len;
        case MIFFT_NORM_INVERSE: return inv ? 1.0 / len : 1.0;
        case MIFFT_NORM_ORTHOGONAL: return 1.0 / std::sqrt(len);
        default: return 1.0;
    }
}

// ------------------------------------------------------------------ builder
static size_t elem_bytes(int prec) { return prec == P_F32 ? 8 : 16; }

// ------------------------------------------------------------ tuning table
namespace {
struct TuningTable {
    std::mutex mu;
    std::map<std::string, std::string> m;
};
TuningTable& tuning_table() {
    static TuningTable t;
    return t;
}
}  // namespace

std::string tuning_lookup(const std::string& key) {
    TuningTable& t = tuning_table();
    std::lock_guard<std::mutex> g(t.mu);
    auto f = t.m.find(key);
    return f == t.m.end() ? std::string() : f->second;
}

void tuning_set(const std::string& key, const std::string& choice) {
    TuningTable& t = tuning_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.m[key] = choice;
}

void tuning_clear() {
    TuningTable& t = tuning_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.m.clear();
}

std::vector<std::pair<std::string, std::string>> tuning_entries() {
    TuningTable& t = tuning_table();
    std::lock_guard<std::mutex> g(t.mu);
    return std::vector<std::pair<std::string, std::string>>(t.m.begin(), t.m.end());
}

static std::string trim(const std::string& x) {
    size_t a = x.find_first_not_of(" \t\r");
    size_t b = x.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
}

static bool valid_choice(const std::string& c) {
    if (c.rfind("reg:", 0) == 0) return find_reg_variant_named(c) != nullptr;
    if (c.rfind("lds:lpw=", 0) == 0) {
        const std::string v = c.substr(8);
        if (v.empty() || v.size() > 4 || v.find_first_not_of("0123456789") != std::string::npos) return false;
        const int lpw = std::stoi(v);
        return lpw >= 1 && lpw <= 1024;
    }
    return false;
}

int tuning_parse(const std::string& text) {
    std::vector<std::pair<std::string, std::string>> got;
    size_t pos = 0;
    int line_no = 0;
    while (pos <= text.size()) {
        size_t nl = text.find('\n', pos);
        if (nl == std::string::npos) nl = text.size();
        const std::string line = trim(text.substr(pos, nl - pos));
        line_no++;
        pos = nl + 1;
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos)
            fail(MIFFT_ERR_INVALID_TUNING, "tuning table line " + std::to_string(line_no) + ": expected 'key = choice'");
        const std::string key = trim(line.substr(0, eq)), choice = trim(line.substr(eq + 1));
        const bool split_key = key.rfind("split|", 0) == 0;
        const bool fold_key = key.rfind("fold|", 0) == 0;
        const bool key_ok = split_key || fold_key || key.rfind("f32|n", 0) == 0 || key.rfind("f64|n", 0) == 0 ||
                            key.rfind("df64|n", 0) == 0;
        // fold|f32|n.. and r.. take the fold factor (0 off, 2, 4 for n); c..
        // (the C2R unfold) and h.. (half-length real axis 0) are on (2) or off (0)
        const bool onoff_key = key.rfind("fold|f32|c", 0) == 0 || key.rfind("fold|f32|h", 0) == 0 ||
                               key.rfind("fold|f32|r", 0) == 0;
        const bool choice_ok = split_key  ? (choice == "2" || choice == "3")
                               : onoff_key ? (choice == "0" || choice == "2")
                               : fold_key ? (choice == "0" || choice == "2" || choice == "4")
                                          : valid_choice(choice);
        if (!key_ok || !choice_ok)
            fail(MIFFT_ERR_INVALID_TUNING, "tuning table line " + std::to_string(line_no) + ": unknown key or choice");
        got.emplace_back(key, choice);
    }
    for (auto& kv : got) tuning_set(kv.first, kv.second);
    return (int)got.size();
}

static std::string tune_key(int prec, uint64_t n, bool ld_lf, bool st_lf, int ldm, int stm, int post, bool pairable,
                            uint64_t lines) {
    static const char* pn[] = {"f32", "f64", "df64"};
    int lb = 0;
    while ((1ull << lb) < lines) lb++;
    char b[160];
    snprintf(b, sizeof(b), "%s|n%llu|lf%d%d|m%d/%d|p%d|pair%d|lines2^%d", pn[prec], (unsigned long long)n, (int)ld_lf,
             (int)st_lf, ldm, stm, post, (int)pairable, lb);
    return b;
}

struct Layout {
    int64_t si, sj, so, sn;
};

struct LineSet {
    uint32_t lines, S1, S2;
};

// Radix schedule of the LDS engine for an n that is {2,3,5,7,11,13}-smooth:
// radix-16 stages first, then the odd primes (3s paired to 9, a 3 and a 5 to 15).
static std::vector<int> lds_schedule(uint64_t n) {
    std::vector<int> r;
    int e2 = 0, e3 = 0, e5 = 0, e7 = 0, e11 = 0, e13 = 0;
    uint64_t m = n;
    while (m % 2 == 0) { m /= 2; e2++; }
    while (m % 3 == 0) { m /= 3; e3++; }
    while (m % 5 == 0) { m /= 5; e5++; }
    while (m % 7 == 0) { m /= 7; e7++; }
    while (m % 11 == 0) { m /= 11; e11++; }
    while (m % 13 == 0) { m /= 13; e13++; }
    while (e2 >= 4) { r.push_back(16); e2 -= 4; }
    if (e2 == 3) r.push_back(8);
    if (e2 == 2) r.push_back(4);
    if (e2 == 1) {
        if (e3 > 0 && (e3 % 2 == 1)) { r.push_back(6); e3--; }
        else r.push_back(2);
    }
    while (e3 >= 2) { r.push_back(9); e3 -= 2; }
    if (e3 == 1) {
        if (e5 > 0) { r.push_back(15); e5--; }
        else r.push_back(3);
    }
    for (int i = 0; i < e5; i++) r.push_back(5);
    for (int i = 0; i < e7; i++) r.push_back(7);
    for (int i = 0; i < e11; i++) r.push_back(11);
    for (int i = 0; i < e13; i++) r.push_back(13);
    return r;
}

// Medium prime radices of the fused Rader convolution
// (FUSED_PRIME_RADICES, src/runtime/axis_plan.rs:2392).
static const int kMediumPrimes[] = {17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
// Fewest butterflies of the largest medium prime per line
// (MIN_MEDIUM_PRIME_BUTTERFLIES, src/runtime/rader_axis.rs:1185-1192).
constexpr uint64_t kMinMediumPrimeButterflies = 36;

// Schedule of the cyclic Rader convolution over p - 1 points when p - 1 has
// prime factors from 17 to 61 besides {2,3,5,7,11,13}
// (medium_prime_rader_schedule, src/runtime/rader_axis.rs:1194-1225): those
// primes' stages, largest first, then the LDS schedule of the smooth rest.
// False when p - 1 is smooth, has a larger prime factor, or holds fewer than
// 36 butterflies of its largest medium prime.
bool medium_rader_schedule(uint64_t p, std::vector<int>* out) {
    if (p < 3) return false;
    const uint64_t l = p - 1;
    if (l < 2 || is_smooth(l)) return false;
    uint64_t rest = l;
    std::vector<int> r;
    for (int i = (int)(sizeof(kMediumPrimes) / sizeof(int)) - 1; i >= 0; i--)
        while (rest % (uint64_t)kMediumPrimes[i] == 0) {
            r.push_back(kMediumPrimes[i]);
            rest /= (uint64_t)kMediumPrimes[i];
        }
    if (r.empty() || l / (uint64_t)r[0] < kMinMediumPrimeButterflies || !is_smooth(rest)) return false;
    if (rest > 1)
        for (int x : lds_schedule(rest)) r.push_back(x);
    if (out) *out = r;
    return true;
}

static Sched make_sched(uint64_t n, uint32_t lpw, bool ld_lf, bool st_lf, const std::vector<int>* radices = nullptr) {
    Sched sc;
    std::memset(&sc, 0, sizeof(sc));
    const std::vector<int> rad = radices ? *radices : lds_schedule(n);
    if (rad.size() > (size_t)kMaxStages) fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "too many stages");
    sc.n = (uint32_t)n;
    sc.ns = (uint32_t)rad.size();
    sc.lpw = lpw;
    sc.ld_lf = ld_lf;
    sc.st_lf = st_lf;
    sc.dn = make_magic((uint32_t)n);
    sc.dlpw = make_magic(lpw);
    uint32_t L = 1;
    uint64_t twoff = 0;
    for (size_t s = 0; s < rad.size(); s++) {
        sc.rad[s] = (uint8_t)rad[s];
        sc.L[s] = L;
        sc.dnr[s] = make_magic((uint32_t)(n / rad[s]));
        sc.dl[s] = make_magic(L);
        sc.twoff[s] = (uint32_t)twoff;
        if (s >= 1) twoff += (uint64_t)(rad[s] - 1) * L;
        L *= rad[s];
    }
    return sc;
}

static std::string sched_str(const Sched& sc) {
    std::string s;
    for (uint32_t i = 0; i < sc.ns; i++) s += (i ? "x" : "") + std::to_string(sc.rad[i]);
    return s.empty() ? "1" : s;
}

// Row pitch (elements) of a workspace copy that a later pass reads down its
// columns: a pass reading narrow row segments (64-byte column pairs) one row
// after another at a pitch that is a multiple of 128 bytes sends a
// workgroup's rows to the same few HBM channels; 64 more bytes per row spread
// them.  Measured on MI355X (4096 x 2048 f32 pair): pads of 64, 192, 320 and
// 576 bytes run 4-5 % faster than none, 128 and 256 bytes no faster.
static uint64_t workspace_pitch(uint64_t elems, size_t es) {
    return (elems * es) % 128 == 0 ? elems + 64 / es : elems;
}

static bool adjacent_lines(const LineSet& ls, const Layout& l) {
    if (l.sn == 1) return false;
    if (ls.S1 > 1) return l.si == 1;
    return ls.S2 > 1 && l.sj == 1;
}

// Estimated cost per point of one four-step pass over lines of length n:
// register-resident kernels stream at close to the copy rate; runtime-
// schedule LDS lines run well when many fit a workgroup (short lines) and
// poorly when only a few do (long f64 lines: narrow strided row segments).
static double pass_cost(int prec, uint64_t n) {
    if (n <= (1u << 30) && !find_reg_variants(prec, (int)n, 1, 1, LD_COMPLEX, ST_COMPLEX, 1).empty()) return 1.0;
    const size_t es = prec == P_F32 ? 8 : 16;
    const size_t line = 2 * (size_t)lds_line_stride((uint32_t)std::min<uint64_t>(n, 1u << 30)) * es;
    const size_t per_wg = kLdsBudget / std::max<size_t>(line, 1);
    return per_wg >= 8 ? 1.5 : per_wg >= 4 ? 2.5 : 6.0;
}

// Best four-step split n = n1 * n2 (both smooth, both on chip under
// `budget`) by estimated cost, ties broken toward balance; 0 if none.
static uint64_t best_split(int prec, uint64_t n, size_t budget, double* cost_out) {
    uint64_t best1 = 0;
    double best = 1e30;
    const size_t es = prec == P_F32 ? 8 : 16;
    for (uint64_t n1 = 2; n1 * n1 <= n * 16 && n1 < n; n1++) {
        if (n % n1) continue;
        const uint64_t n2 = n / n1;
        auto fits = [&](uint64_t len) { return is_smooth(len) && 2 * lds_line_stride((uint32_t)len) * es <= budget; };
        if (!fits(n1) || !fits(n2)) continue;
        const double c = pass_cost(prec, n1) + pass_cost(prec, n2) + 0.05 * std::fabs(std::log((double)n1 / (double)n2));
        if (c < best) {
            best = c;
            best1 = n1;
        }
    }
    if (cost_out) *cost_out = best;
    return best1;
}

// Three-pass split n = n3 * n1 * n2 (each smooth, on chip and with a register
// variant for its pass), as balanced as possible; false if none.  For long
// axes whose two-pass factors are long lines: a 4096-point f64 line fills
// 64 KiB, so only two to four fit a workgroup and its strided side moves 32- to
// 64-byte row segments, while 256-point lines go sixteen to a workgroup.
static bool best_split3(int prec, uint64_t n, size_t budget, uint64_t* o3, uint64_t* o1, uint64_t* o2) {
    const size_t es = prec == P_F32 ? 8 : 16;
    auto fits = [&](uint64_t len, int ld_lf, int st_lf, int post) {
        return is_smooth(len) && 2 * lds_line_stride((uint32_t)len) * es <= budget &&
               !find_reg_variants(prec, (int)len, ld_lf, st_lf, LD_COMPLEX, ST_COMPLEX, post).empty();
    };
    double best = 1e30;
    for (uint64_t a = 2; a * a <= n && a <= 4096; a++) {
        if (n % a || !fits(a, 1, 1, 1)) continue;
        const uint64_t m = n / a;
        for (uint64_t b = 2; b < m && b <= 4096; b++) {
            if (m % b) continue;
            const uint64_t c = m / b;
            if (c > 4096 || !fits(b, 0, 1, 0) || !fits(c, 1, 1, 1)) continue;
            const double hi = (double)std::max(a, std::max(b, c)), lo = (double)std::min(a, std::min(b, c));
            const double cost = std::log(hi / lo);
            if (cost < best) {
                best = cost;
                *o3 = a;
                *o1 = b;
                *o2 = c;
            }
        }
    }
    return best < 1e29;
}

// Whether a long axis takes the three-pass split: by the tuning-table entry
// "split|<prec>|n<n>" ("2" or "3") when present, else when the two-pass
// split's longer factor fills more than 32 KiB per line or the axis has at
// least 2^19 points, except that f32 splits of at most 2048 x 2048 keep two
// passes when the O transforms of the plan hold 2^23 points or more
// together (every pass then has thousands of lines).  Measured on MI355X
// (forward+inverse pairs, f32): one transform, three passes win at
// 2^19-2^21 (1.1-1.4x: the two passes of 2^20 have 1024 lines each, too few
// workgroups for 256 CUs) and 2^23-2^26 (1.0-3.7x), lose 13 % at 2^22 (the
// MI355X table keeps two there); batched, two passes win at 2^23 points
// and more (2^19 x 16 114.6 vs 170.8 us, 2^20 x 8 129.7 vs 161.4, 2^21 x 4
// 135.9 vs 159.6, 2^22 x 2 134.2 vs 157.0) and lose below (2^19 x 8 95.3 vs
// 87.8, 2^20 x 4 112.4 vs 84.9, 2^21 x 2 87.4 vs 85.8); f64 / df64 2^24
// 1.4-1.5x.
static bool prefer_three_pass(int prec, uint64_t n, uint64_t n1, uint64_t n2, uint64_t O) {
    static const char* pn[] = {"f32", "f64", "df64"};
    const std::string forced = tuning_lookup(std::string("split|") + pn[prec] + "|n" + std::to_string(n));
    if (forced == "2") return false;
    if (forced == "3") return true;
    if (prec == P_F32 && std::max(n1, n2) <= 2048 && n * O >= (1ull << 23)) return false;
    const size_t es = prec == P_F32 ? 8 : 16;
    return std::max(n1, n2) * es > 32768 || n >= (1u << 19);
}

struct Builder {
    Plan& P;
    hipStream_t s;
    std::map<std::string, const void*> tw_cache;

    // A plan-owned device allocation; `what` describes a table for
    // mifft_plan_table (empty: not listed).
    void* alloc(size_t bytes, const std::string& what = std::string()) {
        DevBuf b;
        b.bytes = bytes;
        hipError_t e = hipMalloc(&b.p, bytes ? bytes : 16);
        if (e != hipSuccess)
            fail(MIFFT_ERR_ALLOCATION_FAILED, "plan table allocation of " + std::to_string(bytes) + " bytes");
        P.owned.push_back(b);
        if (!what.empty()) P.tables.push_back(TableInfo{b.p, (uint64_t)bytes, what});
        return b.p;
    }
    static const char* prec_name(int arith) { return arith == P_F32 ? "f32" : "f64"; }

    // Stage-twiddle table of a radix schedule, arithmetic precision, generated
    // in FP64 on the device and rounded once.
    const void* stage_twiddles(int arith, const std::vector<int>& rad) {
        std::string key = std::to_string(arith) + ":";
        for (int r : rad) key += std::to_string(r) + "x";
        auto it = tw_cache.find(key);
        if (it != tw_cache.end()) return it->second;
        const uint64_t cnt = stage_twiddle_count(rad.data(), (int)rad.size());
        std::string rs;
        for (size_t i = 0; i < rad.size(); i++) rs += (i ? "," : "") + std::to_string(rad[i]);
        void* t = alloc(std::max<uint64_t>(cnt, 1) * elem_bytes(arith),
                        std::string("stage-twiddles prec=") + prec_name(arith) + " radices=" + rs +
                            " count=" + std::to_string(cnt));
        hip_check(gen_stage_twiddles(arith, t, rad.data(), (int)rad.size(), s), "twiddle generation");
        tw_cache[key] = t;
        return t;
    }

    static std::vector<int> parse_radices(const char* s) {
        std::vector<int> r;
        int v = 0;
        for (const char* c = s;; c++) {
            if (*c >= '0' && *c <= '9') v = v * 10 + (*c - '0');
            else {
                if (v) r.push_back(v);
                v = 0;
                if (!*c) break;
            }
        }
        return r;
    }

    static std::vector<int> sched_radices(const Sched& sc) {
        std::vector<int> r;
        for (uint32_t i = 0; i < sc.ns; i++) r.push_back(sc.rad[i]);
        return r;
    }

    LineArgs base_args(const LineSet& ls, const Layout& in, const Layout& out, bool conj_in,
                       bool conj_out, double scale) {
        LineArgs a;
        std::memset(&a, 0, sizeof(a));
        a.lines = ls.lines;
        a.S1 = ls.S1;
        a.S2 = ls.S2;
        a.d1 = make_magic(ls.S1);
        a.d2 = make_magic(ls.S2);
        a.in_si = in.si; a.in_sj = in.sj; a.in_so = in.so; a.in_sn = in.sn;
        a.out_si = out.si; a.out_sj = out.sj; a.out_so = out.so; a.out_sk = out.sn;
        a.tw = nullptr;
        a.scale = scale;
        a.conj_in = conj_in;
        a.conj_out = conj_out;
        return a;
    }

    // Few lines: the default shape of a pass (the first registered variant,
    // measured on arrays of 64 MiB and more) can leave most CUs idle (2048
    // lines of 512 points in 64 workgroups); then the variant of the same
    // shape and radix schedule with the most workgroups, up to two per CU,
    // runs instead.  The schedule fixes a line's arithmetic, so a line's
    // result does not depend on how many lines its plan has.
    const RegVariant* filling_variant(const RegVariant* rv, uint64_t n, uint64_t lines, bool ld_lf, bool st_lf,
                                      int ldm, int stm, int post, bool pairable) {
        const uint64_t want = 2ull * device_cu_count();
        auto wgs = [&](const RegVariant* v) { return std::min<uint64_t>(want, (lines + v->lines - 1) / v->lines); };
        if (wgs(rv) >= want) return rv;
        const RegVariant* best = rv;
        for (const RegVariant* v : find_reg_variants(P.prec, (int)n, ld_lf, st_lf, ldm, stm, post)) {
            if (v->adj || v->persist || v->fold || (v->vec == 2 && !pairable) || v->lds_bytes > P.lds_budget ||
                std::strcmp(v->radices, rv->radices) != 0)
                continue;
            if (wgs(v) > wgs(best)) best = v;
        }
        return best;
    }

    // One line kernel (register variant when compiled, else the LDS kernel).
    // Returns false if the line does not fit on chip.
    bool line_op(uint64_t n, const LineSet& ls, const Layout& in, const Layout& out, int src,
                 uint64_t src_off, int dst, uint64_t dst_off, bool conj_in, bool conj_out,
                 double scale, int ldm, int stm, const double2* post_hi = nullptr,
                 const double2* post_lo = nullptr, uint32_t post_shift = 0, uint32_t post_mask = 0) {
        const int post = post_hi != nullptr;
        const bool ld_lf = adjacent_lines(ls, in);
        const bool st_lf = adjacent_lines(ls, out);
        Op op;
        op.src = src;
        op.src_off = src_off;
        op.dst = dst;
        op.dst_off = dst_off;
        op.prec = P.prec;
        op.ldm = ldm;
        op.stm = stm;
        op.post = post;
        // Pairs of adjacent lines (16-byte accesses) need both sides to place line
        // g+1 right after line g with every stride even (16-byte alignment).
        // (Pairing lines j and j + 1 of the four-step passes, whose i index
        // is trivial, measured slower: the pair variants' wider workgroups
        // leave too few of them for these line counts.)
        const bool pairable = ld_lf && st_lf && ldm == LD_COMPLEX && stm == ST_COMPLEX && ls.lines % 2 == 0 &&
                              ls.S1 % 2 == 0 && in.si == 1 && out.si == 1 && in.sj % 2 == 0 &&
                              in.so % 2 == 0 && in.sn % 2 == 0 && out.sj % 2 == 0 && out.so % 2 == 0 &&
                              out.sn % 2 == 0;
        // Adjacent-butterfly variants read and write points j, j + 1 of a line
        // in one 16-byte access: unit point strides, even line length, every
        // line at an even element offset.
        auto even_lines = [&](const Layout& l, uint64_t off_bytes) {
            return l.sn == 1 && off_bytes % 16 == 0 && (ls.S1 <= 1 || l.si % 2 == 0) &&
                   (ls.S2 <= 1 || l.sj % 2 == 0) && l.so % 2 == 0;
        };
        op.adj_ok = P.prec == P_F32 && n % 2 == 0 && !ld_lf && !st_lf && ldm == LD_COMPLEX && stm == ST_COMPLEX &&
                    !post && even_lines(in, src_off) && even_lines(out, dst_off);
        op.n = (uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFu);
        op.ld_lf = ld_lf;
        op.st_lf = st_lf;
        op.pairable = pairable;
        op.tkey = tune_key(P.prec, n, ld_lf, st_lf, ldm, stm, post, pairable, ls.lines);
        // A measured choice from the tuning table wins over the default
        // (batched plans keep the fixed set of batched classes).
        const std::string choice = P.batch_mode ? std::string() : tuning_lookup(op.tkey);
        const RegVariant* rv = nullptr;
        uint32_t forced_lpw = 0;
        bool tuned = false;
        if (choice.rfind("reg:", 0) == 0) {
            const RegVariant* t = find_reg_variant_named(choice);
            if (t && t->prec == P.prec && t->n == (int)n && t->ld_lf == (int)ld_lf && t->st_lf == (int)st_lf &&
                t->ldm == ldm && t->stm == stm && t->post == post && (t->vec == 1 || pairable) &&
                (!t->adj || op.adj_ok)) {
                rv = t;
                tuned = true;
            }
        } else if (choice.rfind("lds:lpw=", 0) == 0) {
            forced_lpw = (uint32_t)std::strtoul(choice.c_str() + 8, nullptr, 10);
            tuned = forced_lpw > 0;
        }
        if (!tuned && P.batch_mode) {
            // batched plans: the first registered variant with a batched class
            // (line pairs first)
            for (int want_vec = pairable ? 2 : 1; want_vec >= 1 && !rv; want_vec--)
                for (const RegVariant* v : find_reg_variants(P.prec, (int)n, ld_lf, st_lf, ldm, stm, post))
                    if (v->vec == want_vec && batch_class_for(v) >= 0) {
                        rv = v;
                        break;
                    }
        } else if (!tuned) {
            rv = pairable ? find_reg_variant(P.prec, (int)n, ld_lf, st_lf, ldm, stm, post, 2) : nullptr;
            if (!rv) rv = find_reg_variant(P.prec, (int)n, ld_lf, st_lf, ldm, stm, post, 1);
            if (rv) rv = filling_variant(rv, n, ls.lines, ld_lf, st_lf, ldm, stm, post, pairable);
        }
        const size_t es = elem_bytes(P.arith);
        const uint64_t lstr = lds_line_stride((uint32_t)std::min<uint64_t>(n, 1u << 30));
        if (rv && P.batch_mode && batch_class_for(rv) < 0) rv = nullptr;
        if (!rv && 2 * lstr * es > P.lds_budget) return false;
        op.a = base_args(ls, in, out, conj_in, conj_out, scale);
        op.a.post_hi = post_hi;
        op.a.post_lo = post_lo;
        op.a.post_shift = post_shift;
        op.a.post_mask = post_mask;
        std::ostringstream d;
        if (rv) {
            op.kind = OP_REG;
            op.rv = rv;
            op.a.tw = stage_twiddles(P.arith, parse_radices(rv->radices));
            d << "reg N=" << n << " " << rv->radices << " lines=" << ls.lines << " tpl=" << rv->tpl
              << " vec=" << rv->vec
              << " lpw=" << rv->lines;
            if (rv->persist) d << " pipe" << rv->persist;
        } else {
            op.kind = OP_LDS;
            uint32_t fit = (uint32_t)(P.lds_budget / (2 * lstr * es));
            uint32_t lpw;
            if (ld_lf || st_lf) lpw = P.arith == P_F32 ? 16 : 8;
            else lpw = (uint32_t)std::max<uint64_t>(1, 4096 / n);
            if (forced_lpw) lpw = forced_lpw;
            lpw = std::max<uint32_t>(1, std::min(lpw, fit));
            op.sc = make_sched(n, lpw, ld_lf, st_lf);
            op.a.tw = stage_twiddles(P.arith, sched_radices(op.sc));
            d << "lds N=" << n << " " << sched_str(op.sc) << " lines=" << ls.lines << " lpw=" << lpw;
        }
        if (ld_lf || st_lf) d << (ld_lf ? " ldLF" : "") << (st_lf ? " stLF" : "");
        if (post) d << " post";
        if (ldm != LD_COMPLEX || stm != ST_COMPLEX) d << " mode=" << ldm << "/" << stm;
        op.desc = d.str();
        P.ops.push_back(op);
        return true;
    }

    // f64 device FFT of a length-m sequence in place (filter spectra).
    void fft_f64_inplace(double2* seq, uint64_t m) {
        mifft_config cc;
        std::memset(&cc, 0, sizeof(cc));
        cc.rank = 1;
        cc.shape[0] = m;
        cc.naxes = 1;
        cc.axes[0] = 0;
        cc.batch = 1;
        cc.direction = MIFFT_FORWARD;
        cc.normalization = MIFFT_NORM_NONE;
        cc.precision = MIFFT_F64;
        cc.tuning = P.cfg.tuning;
        cc.tuning.n_force_rader_axes = 0;
        cc.tuning.n_force_bluestein_axes = 0;
        std::unique_ptr<Plan> child = build_plan(cc, MIFFT_C2C);
        void* ws = nullptr;
        if (child->ws_bytes) hip_check(hipMalloc(&ws, child->ws_bytes), "filter workspace");
        run_plan(*child, s, seq, seq, ws);
        hip_check(hipStreamSynchronize(s), "filter FFT");
        if (ws) (void)hipFree(ws);
    }

    // FFT_m(seq)/m rounded once to the arithmetic precision.
    const void* filter_spectrum(double2* seq, uint64_t m, const std::string& what) {
        fft_f64_inplace(seq, m);
        void* f = alloc(m * elem_bytes(P.arith), what + " prec=" + prec_name(P.arith) + " m=" + std::to_string(m));
        hip_check(round_table(P.arith, f, seq, m, 1.0 / (double)m, s), "filter rounding");
        return f;
    }

    // Lines per workgroup of a fused convolution of length m: the most that
    // fit LDS up to the preferred count, fewer when the axis has too few
    // lines to give every CU two workgroups (17 lines of a 17 x 17 volume ran
    // in one workgroup).
    bool conv_fits(uint64_t m, uint32_t* lpw_out, bool lf, uint64_t lines) {
        const uint32_t want = lf ? (P.arith == P_F32 ? 16u : 8u) : (uint32_t)std::max<uint64_t>(1, 2048 / m);
        for (uint32_t lpw = want; lpw >= 1; lpw = lpw / 2) {
            if (conv_kernel_bytes(P.arith, (uint32_t)m, lpw) <= P.lds_budget) {
                const uint64_t fill = 2ull * device_cu_count();
                while (lpw > 1 && (lines + lpw - 1) / lpw < fill) lpw /= 2;
                *lpw_out = lpw;
                return true;
            }
            if (lpw == 1) break;
        }
        return false;
    }

    void rader_op(uint64_t p, uint64_t m, uint32_t lpw, const LineSet& ls, const Layout& in,
                  const Layout& out, int src, uint64_t src_off, int dst, uint64_t dst_off,
                  bool conj_in, bool conj_out, double scale, const std::vector<int>* radices = nullptr) {
        const uint64_t g = primitive_root(p);
        const uint64_t ginv = mod_pow_u64(g, p - 2, p);
        std::vector<uint32_t> pin(p - 1), pout(p - 1);
        for (uint64_t q = 0; q + 1 < p; q++) {
            pin[q] = (uint32_t)mod_pow_u64(g, q, p);
            pout[q] = (uint32_t)mod_pow_u64(ginv, q, p);
        }
        uint32_t* dpin = (uint32_t*)alloc(4 * (p - 1), "rader-perm-in p=" + std::to_string(p) + " g=" + std::to_string(g));
        uint32_t* dpout =
            (uint32_t*)alloc(4 * (p - 1), "rader-perm-out p=" + std::to_string(p) + " ginv=" + std::to_string(ginv));
        hip_check(hipMemcpyAsync(dpin, pin.data(), 4 * (p - 1), hipMemcpyHostToDevice, s), "perm upload");
        hip_check(hipMemcpyAsync(dpout, pout.data(), 4 * (p - 1), hipMemcpyHostToDevice, s), "perm upload");
        double2* seq = nullptr;
        hip_check(hipMalloc(&seq, m * sizeof(double2)), "rader sequence");
        hip_check(gen_rader_seq(seq, p, m, ginv, s), "rader sequence");
        const void* filt =
            filter_spectrum(seq, m, "rader-filter p=" + std::to_string(p) + " ginv=" + std::to_string(ginv));
        hip_check(hipStreamSynchronize(s), "rader filter");
        (void)hipFree(seq);
        Op op;
        op.kind = OP_RADER;
        op.src = src;
        op.src_off = src_off;
        op.dst = dst;
        op.dst_off = dst_off;
        op.prec = P.prec;
        op.conv.io = base_args(ls, in, out, conj_in, conj_out, scale);
        op.conv.sc = make_sched(m, lpw, adjacent_lines(ls, in), adjacent_lines(ls, out), radices);
        op.conv.io.tw = stage_twiddles(P.arith, sched_radices(op.conv.sc));
        op.conv.filt = filt;
        op.conv.perm_in = dpin;
        op.conv.perm_out = dpout;
        op.conv.n = (uint32_t)p;
        op.conv.m = (uint32_t)m;
        op.conv.dn = make_magic((uint32_t)p);
        std::ostringstream d;
        d << "rader p=" << p << " M=" << m << " " << sched_str(op.conv.sc) << " lines=" << ls.lines
          << " lpw=" << lpw;
        op.desc = d.str();
        P.ops.push_back(op);
    }

    const void* chirp_table(uint64_t n) {
        void* c = alloc(n * elem_bytes(P.arith), std::string("chirp prec=") + prec_name(P.arith) + " n=" + std::to_string(n));
        hip_check(gen_chirp(P.arith, c, n, s), "chirp generation");
        return c;
    }

    const void* bluestein_filter(uint64_t n, uint64_t m) {
        double2* seq = nullptr;
        hip_check(hipMalloc(&seq, m * sizeof(double2)), "bluestein sequence");
        hip_check(gen_bluestein_seq(seq, n, m, s), "bluestein sequence");
        const void* f = filter_spectrum(seq, m, "bluestein-filter n=" + std::to_string(n));
        hip_check(hipStreamSynchronize(s), "bluestein filter");
        (void)hipFree(seq);
        return f;
    }

    void bluestein_fused_op(uint64_t n, uint64_t m, uint32_t lpw, const LineSet& ls,
                            const Layout& in, const Layout& out, int src, uint64_t src_off,
                            int dst, uint64_t dst_off, bool conj_in, bool conj_out, double scale) {
        Op op;
        op.kind = OP_BLUESTEIN;
        op.src = src;
        op.src_off = src_off;
        op.dst = dst;
        op.dst_off = dst_off;
        op.prec = P.prec;
        op.conv.io = base_args(ls, in, out, conj_in, conj_out, scale);
        op.conv.sc = make_sched(m, lpw, adjacent_lines(ls, in), adjacent_lines(ls, out));
        op.conv.io.tw = stage_twiddles(P.arith, sched_radices(op.conv.sc));
        op.conv.filt = bluestein_filter(n, m);
        op.conv.chirp = chirp_table(n);
        op.conv.n = (uint32_t)n;
        op.conv.m = (uint32_t)m;
        std::ostringstream d;
        d << "bluestein N=" << n << " M=" << m << " " << sched_str(op.conv.sc)
          << " lines=" << ls.lines << " lpw=" << lpw;
        op.desc = d.str();
        P.ops.push_back(op);
    }

    // Register-resident Bluestein (reg_conv.hpp) over the power-of-two
    // convolution length M >= 2N - 1, when a compiled kernel has that length:
    // one kernel per axis whose two transforms run in registers.
    bool reg_bluestein_op(uint64_t n, const LineSet& ls, const Layout& in, const Layout& out, int src,
                          uint64_t src_off, int dst, uint64_t dst_off, bool conj_in, bool conj_out, double scale) {
        if (P.prec != P_F32 || n < 2 || n > 4096) return false;
        uint64_t m = 64;
        while (m < 2 * n - 1) m *= 2;
        const RegConvVariant* v = find_reg_bluestein((uint32_t)m);
        if (!v || v->lds_bytes > P.lds_budget) return false;
        Op op;
        op.kind = OP_BLUESTEIN;
        op.rcv = v;
        op.src = src;
        op.src_off = src_off;
        op.dst = dst;
        op.dst_off = dst_off;
        op.prec = P.prec;
        op.conv.io = base_args(ls, in, out, conj_in, conj_out, scale);
        op.conv.io.tw = stage_twiddles(P.arith, parse_radices(v->radices));
        op.conv.filt = bluestein_filter(n, m);
        op.conv.chirp = chirp_table(n);
        op.conv.n = (uint32_t)n;
        op.conv.m = (uint32_t)m;
        std::ostringstream d;
        d << "bluestein-reg N=" << n << " M=" << m << " " << v->radices << " lines=" << ls.lines
          << " tpl=" << v->tpl << " lpw=" << v->lines;
        op.desc = d.str();
        P.ops.push_back(op);
        return true;
    }

    // Convolution length of a multi-pass Bluestein axis: any smooth M >= 2N-1
    // serves (bluestein_axis.rs:1035-1041 takes the smallest); among the smooth
    // lengths up to 6 % above the smallest, the one whose padded transform
    // splits into the cheapest passes (for example 4096 x 520 with a
    // register-resident side instead of 1440 x 1458 through LDS).
    uint64_t multipass_conv_len(uint64_t n, uint64_t m0) const {
        uint64_t best_m = m0;
        double best = 1e30;
        for (uint64_t m = m0; m <= m0 + m0 / 16; m = next_smooth(m + 1)) {
            double c = 0;
            if (!best_split(P.arith, m, P.lds_budget, &c)) continue;
            const double total = c * (double)m;
            if (total < best) {
                best = total;
                best_m = m;
            }
        }
        return best_m;
    }

    // Multi-pass Bluestein for convolutions that do not fit LDS; returns the
    // workspace bytes used from ws_off.
    uint64_t bluestein_multi(uint64_t n, uint64_t m, const LineSet& ls, const Layout& in,
                             const Layout& out, int src, uint64_t src_off, int dst,
                             uint64_t dst_off, uint64_t ws_off, bool conj_in, bool conj_out,
                             double scale) {
        if (P.prec == P_DF64)
            fail(MIFFT_ERR_PRECISION_UNSUPPORTED, "df64 Bluestein beyond the on-chip convolution length");
        const uint64_t region = (uint64_t)ls.lines * m * elem_bytes(P.arith);
        mifft_config cc;
        std::memset(&cc, 0, sizeof(cc));
        cc.rank = 1;
        cc.shape[0] = m;
        cc.naxes = 1;
        cc.batch = ls.lines;
        cc.direction = MIFFT_FORWARD;
        cc.normalization = MIFFT_NORM_NONE;
        cc.precision = P.arith == P_F32 ? MIFFT_F32 : MIFFT_F64;
        cc.tuning = P.cfg.tuning;
        cc.tuning.n_force_rader_axes = 0;
        cc.tuning.n_force_bluestein_axes = 0;
        // In batched mode the padded transforms are batched-mode plans too, so
        // that the batched planner can flatten them into its phases.
        PlanOptions copt;
        copt.lds_budget = P.lds_budget;
        copt.batch_mode = P.batch_mode;
        P.children.push_back(build_plan(cc, MIFFT_C2C, copt));
        Plan* child = P.children.back().get();
        const void* chirp = chirp_table(n);
        const void* filt = bluestein_filter(n, m);
        LineArgs a = base_args(ls, in, out, conj_in, conj_out, scale);
        Op pre;
        pre.kind = OP_CHIRP_PRE;
        pre.src = src;
        pre.src_off = src_off;
        pre.dst = B_WS;
        pre.dst_off = ws_off;
        pre.prec = P.prec;
        pre.a = a;
        pre.a.conj_out = 0;
        pre.a.scale = 1.0;
        pre.chirp = chirp;
        pre.n = (uint32_t)n;
        pre.m = (uint32_t)m;
        pre.desc = "bluestein-pre N=" + std::to_string(n) + " M=" + std::to_string(m);
        P.ops.push_back(pre);
        Op sub;
        sub.kind = OP_SUBPLAN;
        sub.dst = B_WS;
        sub.dst_off = ws_off;
        sub.sub = child;
        sub.sub_ws_off = ws_off + region;
        sub.desc = "subplan FFT M=" + std::to_string(m) + " x" + std::to_string(ls.lines);
        P.ops.push_back(sub);
        Op pw;
        pw.kind = OP_POINTWISE;
        pw.dst = B_WS;
        pw.dst_off = ws_off;
        pw.prec = P.prec;
        pw.filt = filt;
        pw.m = (uint32_t)m;
        pw.count = ls.lines;
        pw.desc = "bluestein-filter M=" + std::to_string(m);
        P.ops.push_back(pw);
        P.ops.push_back(sub);
        Op post;
        post.kind = OP_CHIRP_POST;
        post.src = B_WS;
        post.src_off = ws_off;
        post.dst = dst;
        post.dst_off = dst_off;
        post.prec = P.prec;
        post.a = a;
        post.a.conj_in = 0;
        post.chirp = chirp;
        post.n = (uint32_t)n;
        post.m = (uint32_t)m;
        post.desc = "bluestein-post N=" + std::to_string(n);
        P.ops.push_back(post);
        return region + child->ws_bytes;
    }

    // small_volume_key (src/runtime/small_volume.rs:215-275): an f32
    // transform of every axis of a volume of at most 4096 points whose axis
    // lengths are {2,3,5,7,11,13}-smooth or primes up to direct_max_prime
    // (as direct DFTs, unless Rader is forced on the axis; small_volume.rs:
    // 9-10, 220-245) runs as one workgroup per volume.  Not in batched plans (the batched kernel has no
    // class for it), nor for 2D volumes with both sides of 64 points or more:
    // measured on MI355X, 64x64 runs 1.7x faster as two register-kernel
    // passes, while 16x16 ... 32x128, 4x1024, 16^3 and 12x10x9 run 1.1-2x
    // faster in one workgroup.
    bool small_volume(const mifft_config& cfg, uint64_t elems, bool inverse, double scale) {
        if (P.batch_mode || P.prec != P_F32 || cfg.rank < 2 || cfg.naxes != cfg.rank ||
            !cfg.tuning.fuse_small_volumes || elems > 4096)
            return false;
        if (cfg.rank == 2 && std::min(cfg.shape[0], cfg.shape[1]) >= 64) return false;
        const mifft_tuning& tun = cfg.tuning;
        auto direct_prime_axis = [&](int a) {
            return P.axes[a].kind == MIFFT_AXIS_RADER && cfg.shape[a] <= tun.direct_max_prime &&
                   !rader_forced(tun, a);
        };
        // Volumes whose every axis has at most 32 points run the smooth axes
        // as direct DFTs too: n terms per output against a runtime-schedule
        // stage per radix with its barriers (16x16 pair 14.7 -> 9.1 us,
        // 8x8x8 x8 17.1 -> 10.8 us, 12x10x9 x100 22.8 -> 16.5 us, 32x32 18.7
        // -> 15.4 us; a 32-point axis beside a 64-point one measured slower).
        bool all_short = true;
        for (int a = 0; a < cfg.rank; a++) all_short = all_short && cfg.shape[a] <= 32;
        auto direct_axis = [&](int a) {
            return direct_prime_axis(a) || (all_short && P.axes[a].kind == MIFFT_AXIS_MIXED && cfg.shape[a] > 1);
        };
        for (int a = 0; a < cfg.rank; a++)
            if (P.axes[a].kind != MIFFT_AXIS_MIXED && !direct_axis(a)) return false;
        SmallVolArgs v;
        std::memset(&v, 0, sizeof(v));
        v.elems = (uint32_t)elems;
        v.rank = (uint32_t)cfg.rank;
        std::vector<Sched> scs;
        std::vector<const void*> tabs;
        uint32_t S = 1;
        for (int a = 0; a < cfg.rank; a++) {
            const uint32_t n = (uint32_t)cfg.shape[a];
            Sched sc;
            v.S[a] = S;
            v.dS[a] = make_magic(S);
            v.dL[a] = make_magic((uint32_t)(elems / n));
            if (direct_axis(a)) {
                // no stages (ns = 0): the kernel runs the direct DFT with W_n^t, t < n
                std::memset(&sc, 0, sizeof(sc));
                sc.n = n;
                sc.dn = make_magic(n);
                void* wt = alloc(n * sizeof(cpx<float>),
                                 "roots prec=f32 n=" + std::to_string(n) + " count=" + std::to_string(n));
                hip_check(gen_twiddles(P_F32, wt, n, n, s), "direct twiddles");
                tabs.push_back(wt);
                v.tw_len[a] = n;
                v.direct = 1;
            } else {
                sc = make_sched(n, 1, false, false);
                tabs.push_back(stage_twiddles(P_F32, sched_radices(sc)));
                v.tw_len[a] = (uint32_t)stage_twiddle_count(sched_radices(sc).data(), (int)sc.ns);
            }
            v.tw_total += v.tw_len[a];
            v.buf_elems = std::max<uint32_t>(v.buf_elems, (uint32_t)(elems / n) * lds_line_stride(n));
            scs.push_back(sc);
            S *= n;
        }
        if (small_volume_lds_bytes(v) > P.lds_budget) return false;
        Sched* d = (Sched*)alloc(scs.size() * sizeof(Sched));
        hip_check(hipMemcpyAsync(d, scs.data(), scs.size() * sizeof(Sched), hipMemcpyHostToDevice, s),
                  "small-volume schedules");
        // the axes' stage twiddles back to back (one copy into LDS per volume)
        cpx<float>* tw = (cpx<float>*)alloc(std::max<uint32_t>(v.tw_total, 1) * sizeof(cpx<float>));
        for (int a = 0, o = 0; a < cfg.rank; o += v.tw_len[a], a++)
            if (v.tw_len[a])
                hip_check(hipMemcpyAsync(tw + o, tabs[a], v.tw_len[a] * sizeof(cpx<float>), hipMemcpyDeviceToDevice,
                                         s),
                          "small-volume twiddles");
        v.tw = tw;
        hip_check(hipStreamSynchronize(s), "small-volume schedules");
        v.sc = d;
        v.scale = (float)scale;
        v.conj_in = v.conj_out = inverse ? 1u : 0u;
        Op op;
        op.kind = OP_SMALL_VOLUME;
        op.prec = P_F32;
        op.sv = v;
        op.count = cfg.batch;
        std::ostringstream desc;
        desc << "small-volume";
        for (int a = 0; a < cfg.rank; a++) desc << (a ? "x" : " ") << cfg.shape[a];
        desc << " volumes=" << cfg.batch << " lds=" << small_volume_lds_bytes(v);
        op.desc = desc.str();
        P.ops.push_back(op);
        return true;
    }

    // The direct prime DFT (src/runtime/c2c.rs:245-254, rader_axis.rs:290-320):
    // f32 primes up to direct_max_prime whose Rader convolution is not the
    // cyclic one of length p - 1 (p - 1 not smooth), unless Rader is forced
    // on the axis; not in batched plans (no batched class).
    int cur_axis = -1;  // axis being planned, for the forced-axis tunings
    static bool rader_forced(const mifft_tuning& t, int axis) {
        for (int i = 0; i < t.n_force_rader_axes; i++)
            if (t.force_rader_axes[i] == axis) return true;
        return false;
    }
    // Whether the Rader axis of the prime p convolves cyclically over p - 1
    // points with medium prime radices (f32, the fused kernel fits on chip;
    // rader_axis_convolution, src/runtime/rader_axis.rs:1227-1250).
    bool medium_rader(uint64_t p, const LineSet& ls, const Layout& lay, std::vector<int>* radices,
                      uint32_t* lpw) {
        const uint64_t linear = 2 * (p - 1) - 1;
        if (!(P.prec == P_F32 && medium_rader_schedule(p, radices) &&
              std::max<uint64_t>(linear, p - 1) >= P.cfg.tuning.fused_min_convolution_length &&
              conv_fits(p - 1, lpw, adjacent_lines(ls, lay), ls.lines)))
            return false;
        // Contiguous lines of 1024 to about 2400 points: two per tile where
        // two tiles still share a CU (the 512-thread medium-prime kernel
        // leaves most threads idle in the largest prime's stage of one
        // line: N = 1381 over 128 MiB 1.21 -> 0.72 ms per pair; N = 613,
        // three lines per tile, measured 12 % slower at six).
        if (*lpw == 1 && p - 1 >= 1024 && !adjacent_lines(ls, lay) && ls.lines >= 4ull * device_cu_count() &&
            conv_kernel_bytes(P.arith, (uint32_t)(p - 1), 2) <= kLdsBudget / 2)
            *lpw = 2;
        return true;
    }
    bool direct_prime_op(uint64_t p, const LineSet& ls, const Layout& lay, int src, uint64_t src_off, int dst,
                         uint64_t dst_off, bool inverse, double scale) {
        const mifft_tuning& t = P.cfg.tuning;
        if (P.batch_mode || P.prec != P_F32 || p < 3 || p > t.direct_max_prime || rader_conv_len(p) == p - 1)
            return false;
        {
            // a cyclic fused convolution takes priority (c2c.rs:244-253)
            std::vector<int> med;
            uint32_t ml = 0;
            if (medium_rader(p, ls, lay, &med, &ml)) return false;
        }
        for (int i = 0; i < t.n_force_rader_axes; i++)
            if (t.force_rader_axes[i] == cur_axis) return false;
        const bool lf = adjacent_lines(ls, lay);
        const uint32_t tasks = (uint32_t)((p - 1) / 2 + 1);
        uint32_t lpw = std::max<uint32_t>(1, 256 / tasks);
        if (lf) lpw = std::max<uint32_t>(lpw, 8);
        lpw = std::min<uint32_t>(lpw, std::max<uint32_t>(1, ls.lines / 256));  // keep 256 workgroups
        if (direct_prime_bytes((uint32_t)p, lpw) > P.lds_budget) return false;
        void* wt = alloc(p * sizeof(cpx<float>), "roots prec=f32 n=" + std::to_string(p) + " count=" + std::to_string(p));
        hip_check(gen_twiddles(P_F32, wt, p, p, s), "direct twiddles");
        Op op;
        op.kind = OP_DIRECT;
        op.src = src;
        op.src_off = src_off;
        op.dst = dst;
        op.dst_off = dst_off;
        op.prec = P_F32;
        op.a = base_args(ls, lay, lay, inverse, inverse, scale);
        op.n = (uint32_t)p;
        op.m = lpw;
        op.ld_lf = op.st_lf = lf;
        op.filt = wt;
        op.desc = "direct N=" + std::to_string(p) + " lines=" + std::to_string(ls.lines) + " lpw=" +
                  std::to_string(lpw) + (lf ? " LF" : "");
        P.ops.push_back(op);
        return true;
    }

    // Long axis as one Stockham pass per radix over HBM (fuse_long_axes off,
    // or no on-chip split): pass 1 reads the source, the middle passes
    // ping-pong between two workspace copies of the array, the last writes
    // the destination, so no pass reads what it writes.  Returns the
    // workspace bytes used beyond ws_off.
    uint64_t stockham_passes(uint64_t n, uint64_t S, uint64_t O, int src, uint64_t src_off, int dst,
                             uint64_t dst_off, uint64_t ws_off, bool inverse, double scale) {
        if (n > 0xFFFFFFFFull) fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "length beyond 32 bits");
        const std::vector<int> rad = lds_schedule(n);
        const LineSet ls{(uint32_t)(S * O), (uint32_t)S, 1};
        const Layout lay{1, 0, (int64_t)(S * n), (int64_t)S};
        const bool lf = adjacent_lines(ls, lay);
        const void* tw = stage_twiddles(P.arith, rad);
        const uint64_t arr = S * O * n * elem_bytes(P.prec);
        const int ns = (int)rad.size();
        uint32_t L = 1, twoff = 0;
        for (int k = 0; k < ns; k++) {
            const bool first = k == 0, last = k + 1 == ns;
            // ws regions: A at ws_off, B at ws_off + arr; pass k writes A for even k, B for odd k
            const int from = first ? src : B_WS;
            const uint64_t from_off = first ? src_off : ws_off + ((k - 1) % 2 ? arr : 0);
            const int to = last ? dst : B_WS;
            const uint64_t to_off = last ? dst_off : ws_off + (k % 2 ? arr : 0);
            Op op;
            op.kind = OP_STOCKHAM;
            op.src = from;
            op.src_off = from_off;
            op.dst = to;
            op.dst_off = to_off;
            op.prec = P.prec;
            op.a = base_args(ls, lay, lay, first && inverse, last && inverse, last ? scale : 1.0);
            op.a.tw = tw;
            op.st = StockhamStage{(uint32_t)n, (uint32_t)rad[k], L, twoff, lf ? 1u : 0u, make_magic(L)};
            op.desc = "stockham N=" + std::to_string(n) + " R=" + std::to_string(rad[k]) + " L=" + std::to_string(L) +
                      " lines=" + std::to_string(ls.lines);
            P.ops.push_back(op);
            if (k >= 1) twoff += (uint32_t)((rad[k] - 1) * L);
            L *= (uint32_t)rad[k];
        }
        return ns == 1 ? 0 : (ns == 2 ? arr : 2 * arr);
    }

    // Four-step twiddle table W_n^t = hi[t >> sh] * lo[t & mask], t < n.
    struct PostTable {
        const double2* hi;
        const double2* lo;
        uint32_t sh, mask;
    };
    PostTable post_table(uint64_t n) {
        uint32_t sh = 0;
        while ((1ull << (2 * sh)) < n) sh++;
        const uint64_t lo_n = 1ull << sh;
        const uint64_t hi_n = (n + lo_n - 1) / lo_n;
        double2* lo = (double2*)alloc(lo_n * sizeof(double2), "roots-strided prec=f64 n=" + std::to_string(n) +
                                                                   " count=" + std::to_string(lo_n) + " mult=1");
        double2* hi = (double2*)alloc(hi_n * sizeof(double2), "roots-strided prec=f64 n=" + std::to_string(n) +
                                                                   " count=" + std::to_string(hi_n) +
                                                                   " mult=" + std::to_string(lo_n));
        hip_check(gen_twiddles_f64_strided(lo, n, lo_n, 1, s), "post twiddles");
        hip_check(gen_twiddles_f64_strided(hi, n, hi_n, lo_n, s), "post twiddles");
        return PostTable{hi, lo, sh, (uint32_t)(lo_n - 1)};
    }

    // The axis-0 pass of a fold (fold on store) or unfold (Hermitian load,
    // unfold on load): one tile of F rows fold_in / fold_out elements apart,
    // W_{n1}^t from FP64 tables; the tuning table's "...|fold<F>" key picks
    // among the compiled variants fv.
    void fold_row_op(uint64_t n0, uint64_t n1, uint64_t F, int ldm, int stm, const LineSet& ls,
                     const Layout& lay_in, const Layout& lay_out, uint64_t fold_in, uint64_t fold_out, int src,
                     int dst, bool conj, double scale, const std::vector<const RegVariant*>& fv) {
        Op op;
        op.kind = OP_REG;
        op.fold = true;
        op.src = src;
        op.dst = dst;
        op.prec = P.prec;
        op.n = (uint32_t)n0;
        op.ldm = ldm;
        op.stm = stm;
        op.tkey = tune_key(P.prec, n0, false, false, ldm, stm, 0, false, ls.lines) + "|fold" + std::to_string(F);
        // adjacent-butterfly variants: unit point strides, even row pitches
        // and fold offsets (16-byte accesses)
        op.adj_ok = ldm == LD_COMPLEX && stm == ST_COMPLEX && n0 % 2 == 0 && lay_in.sn == 1 && lay_out.sn == 1 &&
                    lay_in.sj % 2 == 0 && lay_in.so % 2 == 0 && lay_out.sj % 2 == 0 && lay_out.so % 2 == 0 &&
                    fold_in % 2 == 0 && fold_out % 2 == 0;
        const RegVariant* rv = find_reg_variant_named(tuning_lookup(op.tkey));
        if (!rv || rv->fold != (int)F || rv->n != (int)n0 || rv->prec != P.prec || rv->ldm != ldm ||
            rv->stm != stm || (rv->adj && !op.adj_ok))
            rv = fv[0];
        op.rv = rv;
        op.a = base_args(ls, lay_in, lay_out, conj, conj, scale);
        op.a.fold_in = (uint32_t)fold_in;
        op.a.fold_out = (uint32_t)fold_out;
        const PostTable pt = post_table(n1);
        op.a.post_hi = pt.hi;
        op.a.post_lo = pt.lo;
        op.a.post_shift = pt.sh;
        op.a.post_mask = pt.mask;
        op.a.tw = stage_twiddles(P.arith, parse_radices(rv->radices));
        std::ostringstream d;
        d << "reg N=" << n0 << " " << rv->radices << " lines=" << ls.lines << " tpl=" << rv->tpl << " vec=" << F
          << " lpw=" << rv->lines;
        if (rv->persist) d << " pipe" << rv->persist;
        d << (ldm == LD_HERMITIAN ? " unfold" : " fold") << F << " M=" << n1;
        if (ldm != LD_COMPLEX || stm != ST_COMPLEX) d << " mode=" << ldm << "/" << stm;
        op.desc = d.str();
        P.ops.push_back(op);
    }

    // C2R of [n0, M] (M = 2H): pass A runs H-point inverse transforms over the
    // packed columns of the even and odd rows of axis 1 (row 2k + c of part
    // c) into halves (row c H + n, workspace); pass B is the C2R axis-0 pass
    // over row pairs (n, n + H), whose load forms E_0 + W_M^{-n} E_1 and
    // E_0 - W_M^{-n} E_1 per packed bin (axis 1's last radix-2
    // decimation-in-time stage) before the Hermitian extension:
    //   x[n + c H] = sum_k X[2k] W_H^{-nk} + (-1)^c W_M^{-n} sum_k X[2k+1] W_H^{-nk}.
    // Table key fold|f32|c<n0>x<M>: 0 off, 2 on; without an entry on when both
    // kernels are compiled and M is at least 512.
    bool unfold_2d_c2r(uint64_t n0, uint64_t n1, uint64_t O, double scale) {
        if (P.batch_mode || P.prec != P_F32 || n1 % 2 != 0 || n1 < 4) return false;
        const std::string choice = tuning_lookup("fold|f32|c" + std::to_string(n0) + "x" + std::to_string(n1));
        if ((!choice.empty() && choice != "2") || (choice.empty() && n1 < 512)) return false;
        const uint64_t H = n1 / 2, P0 = n0 / 2 + 1;
        const std::vector<const RegVariant*> fv = find_reg_variants(P.prec, (int)n0, 0, 0, LD_HERMITIAN, ST_REAL, 0, 2);
        if (fv.empty() || !find_reg_variant(P.prec, (int)H, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2)) return false;
        if (P0 * 2 * O > 0xFFFFFFFFull || n1 * O > 0xFFFFFFFFull) return false;
        const LineSet lsA{(uint32_t)(P0 * 2 * O), (uint32_t)P0, 2};
        const Layout inA{1, (int64_t)P0, (int64_t)(P0 * n1), (int64_t)(2 * P0)};
        const Layout outA{1, (int64_t)(H * P0), (int64_t)(P0 * n1), (int64_t)P0};
        if (!line_op(H, lsA, inA, outA, B_IN, 0, B_WS, 0, true, true, 1.0, LD_COMPLEX, ST_COMPLEX)) return false;
        const LineSet lsB{(uint32_t)(n1 * O), 1, (uint32_t)H};
        const Layout lay_in{0, (int64_t)P0, (int64_t)(P0 * n1), 1};
        const Layout lay_out{0, (int64_t)n0, (int64_t)(n0 * n1), 1};
        fold_row_op(n0, n1, 2, LD_HERMITIAN, ST_REAL, lsB, lay_in, lay_out, H * P0, H * n0, B_WS, B_OUT, true,
                    scale, fv);
        return true;
    }

    // 2D fold: for axes 0 and 1 of lengths n0 and M = F Q (O arrays of n0 x M
    // points), with y the array after the axis-0 transform,
    //   X[F k + c] = DFT_Q(z_c)[k],
    //   z_c[r] = W_M^{r c} sum_b y[r + b Q] W_F^{b c}   (r < Q, rows of axis 1),
    // the first radix-F decimation-in-frequency stage of axis 1 (F = 2: z_0 =
    // y[r] + y[r + Q], z_1 = W_M^r (y[r] - y[r + Q])).  Pass 1 transforms rows
    // r + b Q (b < F) of axis 0 in one tile and stores z_c at row r + c Q
    // (into the workspace); pass 2 runs Q-point lines over the F parts (F n0 O
    // lines) and stores X[F k + c] at row F k + c (into the output).  Axis 1's
    // tile is then 1/F as deep: 8 columns of 1024 points (64 KiB) fit four
    // workgroups per CU where 16 columns of 2048 points fill a CU alone.
    // Table key fold|f32|n<n0>x<M> (r<n0>x<M> for R2C): 0 off, 2 or 4 the
    // factor; without an entry F = 2 when both kernels are compiled and M is
    // at least 512.  R2C
    // (real load, packed store) folds the packed rows of P0 = n0/2 + 1 bins,
    // and pass 2 runs over P0 columns.
    bool fold_2d(uint64_t n0, uint64_t n1, uint64_t O, bool inverse, double scale, uint64_t* ws_bytes,
                 int ldm = LD_COMPLEX, int stm = ST_COMPLEX) {
        if (P.batch_mode || P.prec != P_F32) return false;
        const bool real = ldm == LD_REAL;
        const std::string choice = tuning_lookup(std::string(real ? "fold|f32|r" : "fold|f32|n") +
                                                 std::to_string(n0) + "x" + std::to_string(n1));
        // without an entry the fold runs from M = 512 up: measured neutral to
        // 12 % slower below (M = 128, 256), 7-18 % faster from 512
        if (choice == "0" || (choice.empty() && n1 < 512)) return false;
        const uint64_t F = choice == "4" ? 4 : 2;
        if (n1 % F != 0 || n1 / F < 2) return false;
        const uint64_t Q = n1 / F;
        const uint64_t pin = n0, pout = stm == ST_PACKED ? n0 / 2 + 1 : n0;  // row pitches in elements
        const std::vector<const RegVariant*> fv = find_reg_variants(P.prec, (int)n0, 0, 0, ldm, stm, 0, (int)F);
        if (fv.empty() || !find_reg_variant(P.prec, (int)Q, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2)) return false;
        if (n1 * O > 0xFFFFFFFFull || pout * F * O > 0xFFFFFFFFull) return false;
        // pass 1: axis 0 over row groups (r + b Q, b < F), fold on store
        const LineSet ls1{(uint32_t)(n1 * O), 1, (uint32_t)Q};
        const Layout lay_in{0, (int64_t)pin, (int64_t)(pin * n1), 1};
        // the workspace rows at a pitch of pw elements (see workspace_pitch)
        const uint64_t pw = workspace_pitch(pout, elem_bytes(P.prec));
        if (pw * n1 * O > 0xFFFFFFFFull) return false;
        const Layout lay_out{0, (int64_t)pw, (int64_t)(pw * n1), 1};
        fold_row_op(n0, n1, F, ldm, stm, ls1, lay_in, lay_out, Q * pin, Q * pw, B_IN, B_WS, inverse, 1.0, fv);
        // pass 2: Q-point lines, column i of part c of array o, to rows F k + c
        const uint64_t C = pout;
        const LineSet ls2{(uint32_t)(C * F * O), (uint32_t)C, (uint32_t)F};
        const Layout in2{1, (int64_t)(Q * pw), (int64_t)(pw * n1), (int64_t)pw};
        const Layout out2{1, (int64_t)C, (int64_t)(C * n1), (int64_t)(F * C)};
        if (!line_op(Q, ls2, in2, out2, B_WS, 0, B_OUT, 0, inverse, inverse, scale, LD_COMPLEX, ST_COMPLEX)) {
            P.ops.pop_back();
            return false;
        }
        *ws_bytes = pw * n1 * O * elem_bytes(P.prec);
        return true;
    }

    // Three-pass split n = n3 * n1 * n2 of a contiguous axis (stride 1, O
    // lines 'n' apart), with q = q1 + n1 q2 + n1 n2 q3 and k = k3 + n3 k2 +
    // n3 n2 k1, since W_n^{qk} = W_n3^{q3 k3} W_n^{(q1 + n1 q2) k3}
    // W_n2^{q2 k2} W_{n1 n2}^{q1 k2} W_n1^{q1 k1}:
    //   pass 1: n3-point DFTs over q3 (stride n1 n2), W_n^{(q1 + n1 q2) k3} on
    //           store, in place into the workspace;
    //   pass 2: n2-point DFTs over q2 (stride n1), W_{n1 n2}^{q1 k2} on store,
    //           in place;
    //   pass 3: n1-point DFTs over q1 (contiguous), stored at k3 + n3 k2 +
    //           n3 n2 k1 with adjacent lines (k3, k3 + 1) adjacent in the output.
    // Every pass moves whole row segments of its short lines.
    bool three_pass(uint64_t n, uint64_t O, int src, uint64_t src_off, int dst, uint64_t dst_off,
                    uint64_t ws_off, bool conj_in, bool conj_out, double scale, uint64_t n3, uint64_t n1,
                    uint64_t n2, uint64_t* ws_used) {
        const uint64_t m = n1 * n2;
        const PostTable tn = post_table(n), tm = post_table(m);
        // pass 1: lines (j = q1 + n1 q2 < m, o < O), points q3 at stride m
        const LineSet l1{(uint32_t)(O * m), 1, (uint32_t)m};
        const Layout in1{1, 1, (int64_t)n, (int64_t)m};
        if (!line_op(n3, l1, in1, in1, src, src_off, B_WS, ws_off, conj_in, false, 1.0, LD_COMPLEX, ST_COMPLEX,
                     tn.hi, tn.lo, tn.sh, tn.mask))
            return false;
        // pass 2: lines (j = q1 < n1, o' = k3 + n3 o at stride m), points q2 at stride n1
        const LineSet l2{(uint32_t)(O * n3 * n1), 1, (uint32_t)n1};
        const Layout in2{1, 1, (int64_t)m, (int64_t)n1};
        if (!line_op(n2, l2, in2, in2, B_WS, ws_off, B_WS, ws_off, false, false, 1.0, LD_COMPLEX, ST_COMPLEX,
                     tm.hi, tm.lo, tm.sh, tm.mask))
            return false;
        // pass 3: lines (i = k3 < n3, j = k2 < n2, o < O), points q1 contiguous
        const LineSet l3{(uint32_t)(O * n3 * n2), (uint32_t)n3, (uint32_t)n2};
        const Layout in3{(int64_t)m, (int64_t)n1, (int64_t)n, 1};
        const Layout out3{1, (int64_t)n3, (int64_t)n, (int64_t)(n3 * n2)};
        if (!line_op(n1, l3, in3, out3, B_WS, ws_off, dst, dst_off, false, conj_out, scale, LD_COMPLEX,
                     ST_COMPLEX))
            return false;
        *ws_used = O * n * elem_bytes(P.prec);
        return true;
    }

    // Four-step split n = n1*n2 of an axis too long for one workgroup: pass 1
    // transforms n1-point columns (stride n2) with W_n^(q2 k1) on store into the
    // workspace, pass 2 transforms n2-point rows and stores transposed
    // (k = k1 + n1 k2).  Returns false if no split fits.
    bool four_step(uint64_t n, uint64_t S, uint64_t O, int src, uint64_t src_off, int dst,
                   uint64_t dst_off, uint64_t ws_off, bool conj_in, bool conj_out, double scale,
                   uint64_t* ws_used) {
        // the split with the cheapest estimated passes (register variants
        // where possible), ties broken toward balance
        const uint64_t best1 = best_split(P.arith, n, P.lds_budget, nullptr);
        if (!best1) return false;
        const uint64_t n1 = best1, n2 = n / n1;
        const uint64_t lines_total = S * O;
        // Batched plans keep two passes (measured: no gain in the batched
        // kernel, which already runs many transforms side by side).
        uint64_t t3 = 0, t1 = 0, t2 = 0;
        if (S == 1 && !P.batch_mode && prefer_three_pass(P.prec, n, n1, n2, O) &&
            best_split3(P.prec, n, P.lds_budget, &t3, &t1, &t2)) {
            const size_t mark = P.ops.size();
            if (three_pass(n, O, src, src_off, dst, dst_off, ws_off, conj_in, conj_out, scale, t3, t1, t2, ws_used))
                return true;
            P.ops.resize(mark);
        }
        const PostTable pt = post_table(n);
        const uint64_t N = n;
        // pass 1: lines (i < S, q2 < n2, o < O), points q1 at stride S*n2
        LineSet l1{(uint32_t)(lines_total * n2), (uint32_t)S, (uint32_t)n2};
        Layout in1{1, (int64_t)S, (int64_t)(S * N), (int64_t)(S * n2)};
        Layout ws1 = in1;
        if (!line_op(n1, l1, in1, ws1, src, src_off, B_WS, ws_off, conj_in, false, 1.0, LD_COMPLEX,
                     ST_COMPLEX, pt.hi, pt.lo, pt.sh, pt.mask))
            return false;
        // pass 2: lines (i < S, k1 < n1, o < O), points q2 at stride S
        LineSet l2{(uint32_t)(lines_total * n1), (uint32_t)S, (uint32_t)n1};
        Layout in2{1, (int64_t)(S * n2), (int64_t)(S * N), (int64_t)S};
        Layout out2{1, (int64_t)S, (int64_t)(S * N), (int64_t)(S * n1)};
        if (!line_op(n2, l2, in2, out2, B_WS, ws_off, dst, dst_off, false, conj_out, scale,
                     LD_COMPLEX, ST_COMPLEX))
            return false;
        *ws_used = lines_total * N * elem_bytes(P.prec);
        return true;
    }

    // All ops of one C2C axis: `n` points at stride S, O outer repetitions.
    // Returns the workspace bytes used beyond ws_off.
    uint64_t axis_ops(uint64_t n, int kind, uint64_t S, uint64_t O, int src, uint64_t src_off,
                      int dst, uint64_t dst_off, uint64_t ws_off, bool inverse, double scale) {
        const LineSet ls{(uint32_t)(S * O), (uint32_t)S, 1};
        const Layout lay{1, 0, (int64_t)(S * n), (int64_t)S};
        const bool lf = adjacent_lines(ls, lay);
        if (kind == MIFFT_AXIS_MIXED) {
            if (line_op(n, ls, lay, lay, src, src_off, dst, dst_off, inverse, inverse, scale,
                        LD_COMPLEX, ST_COMPLEX))
                return 0;
            uint64_t used = 0;
            if (P.cfg.tuning.fuse_long_axes &&
                four_step(n, S, O, src, src_off, dst, dst_off, ws_off, inverse, inverse, scale, &used))
                return used;
            return stockham_passes(n, S, O, src, src_off, dst, dst_off, ws_off, inverse, scale);
        }
        if (kind == MIFFT_AXIS_RADER && direct_prime_op(n, ls, lay, src, src_off, dst, dst_off, inverse, scale))
            return 0;
        if (kind == MIFFT_AXIS_RADER) {
            // p - 1 with prime factors from 17 to 61: the cyclic convolution
            // over p - 1 points with those radices, ahead of a zero-padded or
            // register Bluestein one (rader_prefers_register_bluestein,
            // src/runtime/rader_axis.rs:1256-1268)
            std::vector<int> med;
            uint32_t lpw = 0;
            if (medium_rader(n, ls, lay, &med, &lpw)) {
                rader_op(n, n - 1, lpw, ls, lay, lay, src, src_off, dst, dst_off, inverse, inverse, scale, &med);
                return 0;
            }
        }
        // Bluestein axes, and Rader primes whose convolution would be
        // zero-padded (p - 1 not smooth) unless Rader is forced: the
        // register-resident Bluestein kernel (c2c.rs:206-217,
        // rader_axis.rs:1256-1268), on strided axes too.  A strided line's
        // points are a row pitch apart, but the workgroups of neighbouring
        // lines run together on one XCD and share each row segment through
        // its L2 (1229^2 pair 254 -> 113 us against the fused Rader
        // convolution with two lines per workgroup).
        const bool rader_padded = kind == MIFFT_AXIS_RADER && rader_conv_len(n) != n - 1 &&
                                  !(cur_axis >= 0 && rader_forced(P.cfg.tuning, cur_axis));
        if ((kind == MIFFT_AXIS_BLUESTEIN || rader_padded) &&
            reg_bluestein_op(n, ls, lay, lay, src, src_off, dst, dst_off, inverse, inverse, scale))
            return 0;
        if (kind == MIFFT_AXIS_RADER) {
            const uint64_t m = rader_conv_len(n);
            uint32_t lpw = 0;
            if (conv_fits(m, &lpw, lf, ls.lines)) {
                rader_op(n, m, lpw, ls, lay, lay, src, src_off, dst, dst_off, inverse, inverse, scale);
                return 0;
            }
            // Rader convolutions that do not fit on chip run as Bluestein.
        }
        const uint64_t m = bluestein_conv_len(n);
        uint32_t lpw = 0;
        if (conv_fits(m, &lpw, lf, ls.lines)) {
            bluestein_fused_op(n, m, lpw, ls, lay, lay, src, src_off, dst, dst_off, inverse, inverse, scale);
            return 0;
        }
        return bluestein_multi(n, multipass_conv_len(n, m), ls, lay, lay, src, src_off, dst, dst_off, ws_off,
                               inverse, inverse, scale);
    }
};

// axis_kind_for_len_with_rader_max (src/runtime/axis_policy.rs:72-80)
static int kind_for_len(const mifft_tuning& t, uint64_t n) {
    bool ok = false;
    factor_supported(n, &ok);
    if (ok) return MIFFT_AXIS_MIXED;
    if (is_prime_u64(n) && n <= t.rader_max_prime) return MIFFT_AXIS_RADER;
    return MIFFT_AXIS_BLUESTEIN;
}

static bool axis_forced(const mifft_tuning& t, int axis) {
    for (int i = 0; i < t.n_force_rader_axes; i++)
        if (t.force_rader_axes[i] == axis) return true;
    for (int i = 0; i < t.n_force_bluestein_axes; i++)
        if (t.force_bluestein_axes[i] == axis) return true;
    return false;
}

static int axis_kind_for(const mifft_config& c, int axis) {
    const mifft_tuning& t = c.tuning;
    for (int i = 0; i < t.n_force_rader_axes; i++)
        if (t.force_rader_axes[i] == axis) return MIFFT_AXIS_RADER;
    for (int i = 0; i < t.n_force_bluestein_axes; i++)
        if (t.force_bluestein_axes[i] == axis) return MIFFT_AXIS_BLUESTEIN;
    const uint64_t n = c.shape[axis];
    bool ok = false;
    factor_supported(n, &ok);
    if (ok) return MIFFT_AXIS_MIXED;
    if (is_prime_u64(n) && n <= t.rader_max_prime) return MIFFT_AXIS_RADER;
    return MIFFT_AXIS_BLUESTEIN;
}

// Large-chunk route (the reference's LargeChunkPlan, runtime/large_chunk.rs:5-120,
// large_policy.rs:325): a configuration of more than u32::MAX elements is
// split into chunks of whole batches of at most kChunkElems elements, each an
// ordinary plan run at its 64-bit byte offsets, one after another on the
// caller's stream (sharing one workspace).  Every kernel therefore keeps
// 32-bit element indices within a chunk.
static constexpr uint64_t kChunkElems = 1ull << 31;

static std::unique_ptr<Plan> build_large_chunk_plan(const mifft_config& cfg, mifft_kind kind,
                                                    const PlanOptions& opt) {
    uint64_t per = 1;
    for (int i = 0; i < cfg.rank; i++) per *= cfg.shape[i];
    const uint64_t total = per * cfg.batch;
    const uint64_t chunk_batch = std::max<uint64_t>(1, std::min<uint64_t>(cfg.batch, kChunkElems / per));
    auto P = std::make_unique<Plan>();
    P->lds_budget = opt.lds_budget;
    P->cfg = cfg;
    P->kind = kind;
    P->prec = cfg.precision;
    P->arith = cfg.precision == MIFFT_F32 ? P_F32 : P_F64;
    P->large_chunk = true;
    hip_check(hipGetDevice(&P->device), "hipGetDevice");
    auto child = [&](uint64_t b) -> Plan* {
        mifft_config c = cfg;
        c.batch = b;
        P->children.push_back(build_plan(c, kind, opt));
        return P->children.back().get();
    };
    Plan* full = child(chunk_batch);
    const uint64_t rem = cfg.batch % chunk_batch;
    Plan* tail = rem ? child(rem) : nullptr;
    P->route = full->route;
    P->axes = full->axes;
    P->in_place_kernels = full->in_place_kernels;
    const uint64_t in_pb = full->in_bytes / chunk_batch, out_pb = full->out_bytes / chunk_batch;
    P->in_bytes = in_pb * cfg.batch;
    P->out_bytes = out_pb * cfg.batch;
    P->ws_bytes = std::max(full->ws_bytes, tail ? tail->ws_bytes : 0);
    for (uint64_t b0 = 0, k = 0; b0 < cfg.batch; b0 += chunk_batch, k++) {
        Plan* sub = b0 + chunk_batch <= cfg.batch ? full : tail;
        Op op;
        op.kind = OP_CHUNK;
        op.sub = sub;
        op.src = B_IN;
        op.src_off = b0 * in_pb;
        op.dst = B_OUT;
        op.dst_off = b0 * out_pb;
        op.prec = P->prec;
        op.count = sub->cfg.batch;
        std::ostringstream d;
        d << "large-chunk " << k << " batches " << b0 << ".." << (b0 + sub->cfg.batch) << " of " << cfg.batch;
        op.desc = d.str();
        P->ops.push_back(op);
    }
    (void)total;
    return P;
}

std::unique_ptr<Plan> build_plan(const mifft_config& cfg, mifft_kind kind, const PlanOptions& opt) {
    validate_config(cfg);
    if (kind != MIFFT_C2C && kind != MIFFT_R2C && kind != MIFFT_C2R)
        fail(MIFFT_ERR_INVALID_ENUM, "plan kind out of range");
    if (kind != MIFFT_C2C) {
        // validate_real_config / validate_real_precision (src/runtime/real.rs:38-63, 3151-3174)
        const int want = kind == MIFFT_R2C ? MIFFT_FORWARD : MIFFT_INVERSE;
        if (cfg.direction != want)
            fail(MIFFT_ERR_INVALID_REAL_DIRECTION,
                 kind == MIFFT_R2C ? "r2c expects forward" : "c2r expects inverse");
        bool all = cfg.naxes == cfg.rank;
        for (int i = 0; i < cfg.naxes && all; i++) all = cfg.axes[i] == i;
        if (!all) fail(MIFFT_ERR_UNSUPPORTED_REAL_AXES, "real transforms use every axis in order");
        if (cfg.precision != MIFFT_F32)
            fail(MIFFT_ERR_PRECISION_UNSUPPORTED,
                 cfg.precision == MIFFT_F64 ? "real-f64-not-implemented" : "real-df64-not-implemented");
    }
    if (config_total(cfg) > 0xFFFFFFFFull) {
        if (opt.batch_mode) fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batched launches take at most u32::MAX elements");
        return build_large_chunk_plan(cfg, kind, opt);
    }
    auto P = std::make_unique<Plan>();
    P->lds_budget = opt.lds_budget;
    P->batch_mode = opt.batch_mode;
    P->cfg = cfg;
    P->kind = kind;
    P->prec = cfg.precision;
    P->arith = cfg.precision == MIFFT_F32 ? P_F32 : P_F64;
    hip_check(hipGetDevice(&P->device), "hipGetDevice");
    const uint64_t logical = config_total(cfg) / cfg.batch;
    const uint64_t total = config_total(cfg);
    const bool inverse = cfg.direction == MIFFT_INVERSE;
    const double scale = config_scale(cfg);
    for (int k = 0; k < cfg.naxes; k++) {
        AxisInfo ai;
        ai.axis = cfg.axes[k];
        ai.n = cfg.shape[ai.axis];
        ai.kind = axis_kind_for(cfg, ai.axis);
        bool ok = false;
        ai.factors = factor_supported(ai.n, &ok);
        if (!ok) ai.factors.clear();
        P->axes.push_back(ai);
    }
    // select_route (src/runtime/c2c.rs:7631-7650)
    if (total == 1) P->route = MIFFT_ROUTE_DIRECT_DFT;
    else {
        bool all_mixed = true;
        for (auto& a : P->axes) all_mixed &= a.kind == MIFFT_AXIS_MIXED;
        if (all_mixed) P->route = MIFFT_ROUTE_MIXED_RADIX;
        else if (P->axes.size() == 1) P->route = P->axes[0].kind == MIFFT_AXIS_RADER ? MIFFT_ROUTE_RADER : MIFFT_ROUTE_BLUESTEIN;
        else P->route = MIFFT_ROUTE_AXIS_SEQUENCE;
    }
    hipStream_t s;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "plan stream");
    Builder B{*P, s, {}};
    try {
        const size_t cb = elem_bytes(P->prec);
        if (kind == MIFFT_C2C) {
            P->in_bytes = P->out_bytes = total * cb;
            uint64_t fold_ws = 0;
            if (total == 1) {
                Op op;
                op.kind = OP_SCALE_COPY;
                op.prec = P->prec;
                op.count = 1;
                op.scale = scale;
                op.desc = "direct N=1";
                P->ops.push_back(op);
            } else if (B.small_volume(cfg, logical, inverse, scale)) {
                P->ws_bytes = 0;
            } else if (cfg.naxes == 2 && cfg.axes[0] == 0 && cfg.axes[1] == 1 &&
                       P->axes[0].kind == MIFFT_AXIS_MIXED && P->axes[1].kind == MIFFT_AXIS_MIXED &&
                       B.fold_2d(cfg.shape[0], cfg.shape[1], total / (cfg.shape[0] * cfg.shape[1]), inverse,
                                 scale, &fold_ws)) {
                // the fold pair ping-pongs through one workspace copy of the array
                P->ws_bytes = fold_ws;
            } else {
                // Ping-pong: with several axes, consecutive passes alternate between
                // the output and a workspace copy of the array (the last pass lands
                // in the output), so no pass reads and writes the same buffer;
                // in-place streams run measurably slower on HBM than
                // out-of-place ones.  Scratch of multi-pass axes sits after it.
                const uint64_t arr_bytes = total * cb;
                const bool pp = cfg.naxes >= 2;
                const uint64_t scratch = pp ? arr_bytes : 0;
                uint64_t ws = scratch;
                int src = B_IN;
                for (int k = 0; k < cfg.naxes; k++) {
                    const int axis = cfg.axes[k];
                    uint64_t S = 1;
                    for (int d = 0; d < axis; d++) S *= cfg.shape[d];
                    const uint64_t n = cfg.shape[axis];
                    const uint64_t O = total / (S * n);
                    const bool last = k + 1 == cfg.naxes;
                    const int dst = (!pp || (cfg.naxes - 1 - k) % 2 == 0) ? B_OUT : B_WS;
                    B.cur_axis = axis;
                    ws = std::max(ws, scratch + B.axis_ops(n, P->axes[k].kind, S, O, src, 0, dst, 0,
                                                           scratch, inverse, last ? scale : 1.0));
                    src = dst;
                }
                P->ws_bytes = ws;
            }
        } else {
            // Real transforms over shape [N0, N1, ...]: the packed spectrum is
            // [N0/2+1, N1, ...] with axis 0 contiguous (src/math/mod.rs:444-502).
            const uint64_t N0 = cfg.shape[0];
            const uint64_t P0 = N0 / 2 + 1;
            const uint64_t outer = total / N0;  // lines of axis 0 (all batches)
            const uint64_t packed = P0 * outer;
            if (kind == MIFFT_R2C) {
                P->in_bytes = total * 4;
                P->out_bytes = packed * 8;
            } else {
                P->in_bytes = packed * 8;
                P->out_bytes = total * 4;
            }
            // Axis 0 runs fused with the real load / packed store when it is smooth
            // and fits one workgroup; otherwise (Rader, Bluestein, four-step
            // lengths) as a complex transform of full lines in the workspace,
            // between an element-wise widening / packing (R2C) or Hermitian
            // extension / real part (C2R).
            const bool smooth0 = P->axes[0].kind == MIFFT_AXIS_MIXED;
            const LineSet l0{(uint32_t)outer, 1, 1};
            const Layout real_l{0, 0, (int64_t)N0, 1};
            const Layout pack_l{0, 0, (int64_t)P0, 1};
            const uint64_t full_bytes = total * 8;
            auto elem_op = [&](OpKind k, int src, uint64_t src_off, int dst, uint64_t dst_off, uint64_t count,
                               double sc, const char* what) {
                Op op;
                op.kind = k;
                op.src = src;
                op.src_off = src_off;
                op.dst = dst;
                op.dst_off = dst_off;
                op.prec = P_F32;
                op.n = (uint32_t)N0;
                op.count = count;
                op.scale = sc;
                op.desc = std::string(what) + " N0=" + std::to_string(N0) + " lines=" + std::to_string(outer);
                P->ops.push_back(op);
            };
            // Axis 0 of even length that is not fused runs as N0/2-point complex
            // lines (the real line read as complex pairs) plus one twiddle pass,
            // instead of widening to N0-point complex lines: half the points and
            // two passes fewer (kern_util.hip k_r2c_post / k_c2r_pre).  Not for a
            // forced axis 0; table key fold|f32|h<N0> = 0 turns it off.
            const bool half_real = N0 % 2 == 0 && N0 >= 4 && !axis_forced(cfg.tuning, 0) &&
                                   tuning_lookup("fold|f32|h" + std::to_string(N0)) != "0";
            Builder::PostTable half_tw{};
            auto half_op = [&](OpKind k, int src, uint64_t src_off, int dst, uint64_t dst_off, double sc,
                               const char* what) {
                if (!half_tw.hi) half_tw = B.post_table(N0);
                Op op;
                op.kind = k;
                op.src = src;
                op.src_off = src_off;
                op.dst = dst;
                op.dst_off = dst_off;
                op.prec = P_F32;
                op.n = (uint32_t)(N0 / 2);
                op.count = outer;
                op.scale = sc;
                op.a.post_hi = half_tw.hi;
                op.a.post_lo = half_tw.lo;
                op.a.post_shift = half_tw.sh;
                op.a.post_mask = half_tw.mask;
                op.desc = std::string(what) + " N0=" + std::to_string(N0) + " lines=" + std::to_string(outer);
                P->ops.push_back(op);
            };
            uint64_t ws = 0;
            if (kind == MIFFT_R2C && cfg.rank == 2 && smooth0 && P->axes[1].kind == MIFFT_AXIS_MIXED &&
                B.fold_2d(N0, cfg.shape[1], cfg.batch, false, scale, &ws, LD_REAL, ST_PACKED)) {
                // 2D fold: real rows -> folded packed rows in the workspace -> output
            } else if (kind == MIFFT_R2C) {
                const double s0 = cfg.rank == 1 ? scale : 1.0;
                const bool fused0 = smooth0 && B.line_op(N0, l0, real_l, pack_l, B_IN, 0, B_OUT, 0, false, false,
                                                         s0, LD_REAL, ST_PACKED);
                if (!fused0 && half_real) {
                    // the real lines as N0/2-point complex lines, then their bins to packed spectra
                    B.cur_axis = -1;
                    ws = total * 4 + B.axis_ops(N0 / 2, kind_for_len(cfg.tuning, N0 / 2), 1, outer, B_IN, 0, B_WS, 0,
                                                total * 4, false, 1.0);
                    half_op(OP_R2C_POST, B_WS, 0, B_OUT, 0, s0, "r2c post");
                } else if (!fused0) {
                    elem_op(OP_REAL_WIDEN, B_IN, 0, B_WS, 0, total, 1.0, "real widen");
                    B.cur_axis = 0;
                    ws = full_bytes + B.axis_ops(N0, P->axes[0].kind, 1, outer, B_WS, 0, B_WS, 0, full_bytes,
                                                 false, 1.0);
                    elem_op(OP_PACK, B_WS, 0, B_OUT, 0, outer, s0, "pack");
                }
                uint64_t S = P0;
                for (int axis = 1; axis < cfg.rank; axis++) {
                    const uint64_t n = cfg.shape[axis];
                    const uint64_t O = packed / (S * n);
                    B.cur_axis = axis;
                    ws = std::max(ws, B.axis_ops(n, P->axes[axis].kind, S, O, B_OUT, 0, B_OUT, 0, 0,
                                                 false, axis + 1 == cfg.rank ? scale : 1.0));
                    S *= n;
                }
            } else if (cfg.rank == 2 && smooth0 && P->axes[1].kind == MIFFT_AXIS_MIXED &&
                       B.unfold_2d_c2r(N0, cfg.shape[1], cfg.batch, scale)) {
                // 2D unfold: packed spectrum -> halves in the workspace -> real output
                ws = packed * 8;
            } else {
                int src = B_IN;
                uint64_t src_off = 0;
                const uint64_t pk_bytes = packed * 8;
                if (cfg.rank > 1) {
                    uint64_t S = P0;
                    for (int axis = 1; axis < cfg.rank; axis++) {
                        const uint64_t n = cfg.shape[axis];
                        const uint64_t O = packed / (S * n);
                        B.cur_axis = axis;
                        ws = std::max(ws, B.axis_ops(n, P->axes[axis].kind, S, O, src, src_off, B_WS,
                                                     0, pk_bytes, true, 1.0));
                        src = B_WS;
                        src_off = 0;
                        S *= n;
                    }
                    ws += pk_bytes;
                }
                const bool fused0 = smooth0 && B.line_op(N0, l0, pack_l, real_l, src, src_off, B_OUT, 0, true, true,
                                                         scale, LD_HERMITIAN, ST_REAL);
                if (!fused0 && half_real) {
                    // packed spectra -> N0/2-point complex lines after the packed copy of
                    // the outer passes, then their inverse straight into the real output
                    const uint64_t X = cfg.rank > 1 ? pk_bytes : 0;
                    half_op(OP_C2R_PRE, src, src_off, B_WS, X, 1.0, "c2r pre");
                    B.cur_axis = -1;
                    ws = std::max(ws, X + total * 4 +
                                          B.axis_ops(N0 / 2, kind_for_len(cfg.tuning, N0 / 2), 1, outer, B_WS, X,
                                                     B_OUT, 0, X + total * 4, true, scale));
                } else if (!fused0) {
                    // full complex lines after the packed copy of the outer passes
                    const uint64_t X = cfg.rank > 1 ? pk_bytes : 0;
                    elem_op(OP_HERM_EXTEND, src, src_off, B_WS, X, outer, 1.0, "hermitian extend");
                    B.cur_axis = 0;
                    ws = std::max(ws, X + full_bytes +
                                          B.axis_ops(N0, P->axes[0].kind, 1, outer, B_WS, X, B_WS, X,
                                                     X + full_bytes, true, 1.0));
                    elem_op(OP_REAL_PART, B_WS, X, B_OUT, 0, total, scale, "real part");
                }
            }
            P->ws_bytes = ws;
        }
        hip_check(hipStreamSynchronize(s), "plan tables");
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        throw;
    }
    hip_check(hipStreamDestroy(s), "plan stream");
    // In-place execution is safe when every op of the plan reads its whole
    // line set before writing it (single-op axes) or goes through the
    // workspace (four-step, multi-pass Bluestein); real plans change size.
    P->in_place_kernels = kind == MIFFT_C2C;
    return P;
}

static void run_op(const Op& op, hipStream_t s, const void* in, void* out, void* ws) {
    auto ptr = [&](int ref, uint64_t off) -> char* {
        char* b = ref == B_IN ? (char*)in : ref == B_OUT ? (char*)out : (char*)ws;
        return b + off;
    };
    switch (op.kind) {
        case OP_REG: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_reg(op.rv, a, s), "register line kernel");
            break;
        }
        case OP_LDS: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_lds(op.prec, op.ldm, op.stm, op.post, a, op.sc, s), "LDS line kernel");
            break;
        }
        case OP_RADER:
        case OP_BLUESTEIN: {
            ConvArgs c = op.conv;
            c.io.in = ptr(op.src, op.src_off);
            c.io.out = ptr(op.dst, op.dst_off);
            hip_check(op.kind == OP_RADER ? launch_rader(op.prec, c, s)
                      : op.rcv            ? launch_reg_bluestein(op.rcv, c, s)
                                          : launch_bluestein(op.prec, c, s),
                      "convolution kernel");
            break;
        }
        case OP_STOCKHAM: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_stockham_pass(op.prec, a, op.st, s), "Stockham pass");
            break;
        }
        case OP_DIRECT: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_direct_prime(a, op.n, op.m, op.ld_lf, op.filt, s), "direct prime kernel");
            break;
        }
        case OP_SMALL_VOLUME: {
            SmallVolArgs v = op.sv;
            v.in = ptr(op.src, op.src_off);
            v.out = ptr(op.dst, op.dst_off);
            hip_check(launch_small_volume(v, (uint32_t)op.count, s), "small-volume kernel");
            break;
        }
        case OP_CHIRP_PRE: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_chirp_pre(op.prec, a, op.chirp, op.n, op.m, s), "chirp pre");
            break;
        }
        case OP_POINTWISE:
            hip_check(launch_pointwise(op.prec, ptr(op.dst, op.dst_off), op.filt, op.m, op.count, s),
                      "pointwise");
            break;
        case OP_CHIRP_POST: {
            LineArgs a = op.a;
            a.in = ptr(op.src, op.src_off);
            a.out = ptr(op.dst, op.dst_off);
            hip_check(launch_chirp_post(op.prec, a, op.chirp, op.n, op.m, s), "chirp post");
            break;
        }
        case OP_SUBPLAN: {
            char* b = ptr(op.dst, op.dst_off);
            run_plan(*op.sub, s, b, b, (char*)ws + op.sub_ws_off);
            break;
        }
        case OP_CHUNK:
            run_plan(*op.sub, s, ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), ws);
            break;
        case OP_SCALE_COPY:
            hip_check(launch_scale_copy(op.prec, ptr(op.src, op.src_off), ptr(op.dst, op.dst_off),
                                        op.count, op.scale, s),
                      "scale copy");
            break;
        case OP_REAL_WIDEN:
            hip_check(launch_real_widen(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.count, s), "real widen");
            break;
        case OP_PACK:
            hip_check(launch_pack(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.n, op.count, op.scale, s),
                      "pack");
            break;
        case OP_HERM_EXTEND:
            hip_check(launch_herm_extend(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.n, op.count, s),
                      "hermitian extend");
            break;
        case OP_REAL_PART:
            hip_check(launch_real_part(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.count, op.scale, s),
                      "real part");
            break;
        case OP_R2C_POST:
            hip_check(launch_r2c_post(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.n, op.count, op.scale,
                                      op.a.post_hi, op.a.post_lo, op.a.post_shift, op.a.post_mask, s),
                      "r2c post");
            break;
        case OP_C2R_PRE:
            hip_check(launch_c2r_pre(ptr(op.src, op.src_off), ptr(op.dst, op.dst_off), op.n, op.count, op.a.post_hi,
                                     op.a.post_lo, op.a.post_shift, op.a.post_mask, s),
                      "c2r pre");
            break;
    }
}

void run_plan(const Plan& p, hipStream_t s, const void* in, void* out, void* ws) {
    for (const Op& op : p.ops) run_op(op, s, in, out, ws);
}

int plan_launch_count(const Plan& p) {
    int c = 0;
    for (const Op& op : p.ops) c += (op.kind == OP_SUBPLAN || op.kind == OP_CHUNK) ? plan_launch_count(*op.sub) : 1;
    return c;
}

static bool launch_at(const Plan& p, hipStream_t s, bool run, int* index, const void* in, void* out, void* ws,
                      std::string* desc) {
    for (const Op& op : p.ops) {
        if (op.kind == OP_SUBPLAN) {
            char* b = (op.dst == B_IN ? (char*)in : op.dst == B_OUT ? (char*)out : (char*)ws) + op.dst_off;
            if (launch_at(*op.sub, s, run, index, b, b, (char*)ws + op.sub_ws_off, desc)) return true;
            continue;
        }
        if (op.kind == OP_CHUNK) {
            if (launch_at(*op.sub, s, run, index, (const char*)in + op.src_off, (char*)out + op.dst_off, ws, desc))
                return true;
            continue;
        }
        if (*index == 0) {
            if (desc) *desc = op.desc;
            if (run) run_op(op, s, in, out, ws);
            return true;
        }
        (*index)--;
    }
    return false;
}

bool run_plan_launch(const Plan& p, hipStream_t s, bool run, int index, const void* in, void* out, void* ws,
                     std::string* desc) {
    if (index < 0) return false;
    int i = index;
    return launch_at(p, s, run, &i, in, out, ws, desc);
}

static const Op* op_at(const Plan& p, int* index) {
    for (const Op& op : p.ops) {
        if (op.kind == OP_SUBPLAN || op.kind == OP_CHUNK) {
            if (const Op* f = op_at(*op.sub, index)) return f;
            continue;
        }
        if (*index == 0) return &op;
        (*index)--;
    }
    return nullptr;
}

void plan_tables(const Plan& p, std::vector<const TableInfo*>* out) {
    for (const TableInfo& t : p.tables) out->push_back(&t);
    for (const auto& c : p.children) plan_tables(*c, out);
}

const Op* plan_launch_op(const Plan& p, int index) {
    if (index < 0) return nullptr;
    int i = index;
    return op_at(p, &i);
}

std::string describe_plan(const Plan& p) {
    std::ostringstream d;
    bool first = true;
    for (const Op& op : p.ops) {
        if (!first) d << "; ";
        first = false;
        d << op.desc;
    }
    return d.str();
}

// FftPlan::diagnostics (src/plan.rs:938-1004, src/diagnostics.rs:8-140) as
// a JSON document: route summary, one stage per launch (child plans
// expanded), buffer requirements (input, output, workspace, helper tables) and
// the active tuning.  A plan that exists has no blockers.
static const char* op_kind_name(OpKind k) {
    switch (k) {
        case OP_REG: return "register-lines";
        case OP_LDS: return "lds-lines";
        case OP_RADER: return "rader";
        case OP_BLUESTEIN: return "bluestein";
        case OP_CHIRP_PRE: return "bluestein-chirp-pre";
        case OP_POINTWISE: return "bluestein-filter";
        case OP_CHIRP_POST: return "bluestein-chirp-post";
        case OP_SUBPLAN: return "subplan";
        case OP_CHUNK: return "large-chunk";
        case OP_SCALE_COPY: return "direct-dft";
        case OP_REAL_WIDEN: return "real-widen";
        case OP_PACK: return "real-pack";
        case OP_HERM_EXTEND: return "hermitian-extend";
        case OP_REAL_PART: return "real-part";
        case OP_R2C_POST: return "r2c-post";
        case OP_C2R_PRE: return "c2r-pre";
        case OP_SMALL_VOLUME: return "small-volume";
        case OP_DIRECT: return "direct-prime";
        case OP_STOCKHAM: return "stockham-pass";
    }
    return "unknown";
}

static std::string json_str(const std::string& v) {
    std::string o = "\"";
    for (char c : v) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) continue;
        o += c;
    }
    return o + "\"";
}

static void diag_stages(const Plan& p, const std::string& route, std::ostringstream& d, bool& first) {
    for (const Op& op : p.ops) {
        if ((op.kind == OP_SUBPLAN || op.kind == OP_CHUNK) && op.sub) {
            diag_stages(*op.sub, route, d, first);
            continue;
        }
        // kind: the reference's stage kind (every launch here is a kernel);
        // algorithm: which kernel family runs it
        d << (first ? "" : ",") << "{\"label\":" << json_str(op.desc) << ",\"kind\":\"kernel\",\"algorithm\":\""
          << op_kind_name(op.kind) << "\",\"route\":" << json_str(route) << ",\"required_bytes\":null}";
        first = false;
    }
}

static uint64_t table_bytes(const Plan& p) {
    uint64_t b = 0;
    for (const DevBuf& o : p.owned) b += o.bytes;
    for (const auto& c : p.children) b += table_bytes(*c);
    return b;
}

std::string plan_diagnostics_json(const Plan& p, const std::vector<std::string>& blockers) {
    static const char* routes[] = {"direct-dft", "mixed-radix", "rader", "bluestein", "axis-sequence"};
    static const char* kinds[] = {"c2c", "r2c", "c2r"};
    static const char* cfmt[] = {"complex-f32", "complex-f64", "complex-df64"};
    const std::string route = (p.route >= 0 && p.route <= 4) ? routes[p.route] : "unknown";
    const std::string cf = cfmt[p.prec];
    const std::string in_fmt = p.kind == MIFFT_R2C ? "real-f32" : p.kind == MIFFT_C2R ? "packed-complex-f32" : cf;
    const std::string out_fmt = p.kind == MIFFT_R2C ? "packed-complex-f32" : p.kind == MIFFT_C2R ? "real-f32" : cf;
    std::ostringstream d;
    d << "{\"route\":{\"transform\":\"" << kinds[p.kind] << "\",\"route\":" << json_str(route)
      << (p.large_chunk ? ",\"large_route_mode\":\"large-chunk\",\"execution_kind\":\"large-chunk\","
                          "\"reason_codes\":[\"exceeds-u32-index-space\"],\"attempted_routes\":[\"direct\",\"batch-chunk\"]},"
                        : ",\"large_route_mode\":\"normal\",\"execution_kind\":\"normal\","
                          "\"reason_codes\":[\"within-bindings\"],\"attempted_routes\":[\"direct\"]},");
    d << "\"stages\":[";
    bool first = true;
    diag_stages(p, route, d, first);
    d << "],\"blockers\":[";
    for (size_t i = 0; i < blockers.size(); i++) d << (i ? "," : "") << blockers[i];
    d << "],";
    d << "\"buffer_requirements\":[{\"role\":\"input\",\"required_bytes\":" << p.in_bytes << ",\"format\":\""
      << in_fmt << "\"},{\"role\":\"output\",\"required_bytes\":" << p.out_bytes << ",\"format\":\"" << out_fmt
      << "\"}";
    if (p.ws_bytes)
        d << ",{\"role\":\"workspace\",\"required_bytes\":" << p.ws_bytes << ",\"format\":\"" << cf << "\"}";
    const uint64_t tb = table_bytes(p);
    if (tb) d << ",{\"role\":\"helper:twiddle-luts-total\",\"required_bytes\":" << tb << ",\"format\":\"" << cf << "\"}";
    d << "],";
    d << "\"active_tuning\":" << tuning_json(p.cfg.tuning) << ",";
    d << "\"device_limits\":{\"max_workgroup_lds_bytes\":" << kLdsBudget << ",\"device\":" << p.device << "}}";
    return d.str();
}

// ------------------------------------------------------------- autotuner
int autotune_sequence(const mifft_config* cfgs, const mifft_kind* kinds, int count, int reps, hipStream_t s,
                      std::string* report) {
    if (reps < 1 || reps > 10000) fail(MIFFT_ERR_INVALID_TUNING, "autotune repetitions must be 1..10000");
    if (count < 1 || count > 16) fail(MIFFT_ERR_INVALID_TUNING, "autotune takes 1..16 plans");
    std::vector<std::unique_ptr<Plan>> plans;
    for (int i = 0; i < count; i++) plans.push_back(build_plan(cfgs[i], kinds[i]));
    // holder of candidate tables and scratch buffers
    Plan T;
    T.cfg = plans[0]->cfg;
    T.prec = plans[0]->prec;
    T.arith = plans[0]->arith;
    T.lds_budget = plans[0]->lds_budget;
    auto alloc = [&](size_t bytes) -> void* {
        DevBuf b;
        b.bytes = std::max<size_t>(bytes, 256);
        hip_check(hipMalloc(&b.p, b.bytes), "autotune buffer");
        T.owned.push_back(b);
        hip_check(hipMemsetAsync(b.p, 0, b.bytes, s), "autotune buffer");
        return b.p;
    };
    // The sequence runs as a caller runs it: in place on one array when every
    // plan allows it (a forward+inverse pair), otherwise each plan reading the
    // previous plan's output (R2C then C2R).  Arrays rotate over sets larger
    // than the 256 MB MALL so each execution streams from HBM; each plan has
    // one workspace, as it executes with its own.
    bool in_place = true;
    size_t arr = 256;
    for (auto& p : plans) {
        in_place &= p->in_place_kernels;
        arr = std::max({arr, (size_t)p->in_bytes, (size_t)p->out_bytes});
    }
    const int nsets = (int)std::min<size_t>(8, std::max<size_t>(1, ((size_t)512 << 20) / (2 * arr) + 1));
    std::vector<void*> bufa, bufb, wsp;
    for (int k = 0; k < nsets; k++) {
        bufa.push_back(alloc(arr));
        bufb.push_back(in_place ? nullptr : alloc(arr));
    }
    for (auto& p : plans) wsp.push_back(alloc(p->ws_bytes));
    Builder TB{T, s, {}};
    std::vector<std::vector<Op>> ops;
    for (auto& p : plans) ops.push_back(p->ops);
    auto run_all = [&](int k) {
        void* x = bufa[k % nsets];
        void* y = in_place ? x : bufb[k % nsets];
        for (size_t pi = 0; pi < ops.size(); pi++) {
            for (const Op& o : ops[pi]) run_op(o, s, x, y, wsp[pi]);
            if (!in_place) std::swap(x, y);
        }
    };
    // distinct tunable op shapes, in launch order
    std::vector<std::string> keys;
    for (auto& p : plans)
        for (const Op& o : p->ops)
            if ((o.kind == OP_REG || o.kind == OP_LDS) && !o.tkey.empty() &&
                std::find(keys.begin(), keys.end(), o.tkey) == keys.end())
                keys.push_back(o.tkey);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "autotune event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        fail(MIFFT_ERR_HIP, "autotune event");
    }
    // Coordinate descent over the shapes: every candidate of a shape is
    // applied to all ops of that shape in every plan and the whole sequence is
    // timed (so each pass sees the cache state its predecessor leaves: a pass
    // reading an array its predecessor just wrote partly hits the MALL); the
    // winner stays for the shapes tuned after it.
    int tuned = 0;
    try {
        for (const std::string& key : keys) {
            const Op* ref = nullptr;
            for (auto& p : plans)
                for (const Op& o : p->ops)
                    if (!ref && o.tkey == key) ref = &o;
            // candidate (name, op template: kind, rv or schedule, tables)
            std::vector<std::pair<std::string, Op>> cands;
            // a variant applies to every op of the key: adjacent-butterfly
            // forms only where each of them allows 16-byte accesses
            bool adj_ok = true;
            for (auto& p : plans)
                for (const Op& o : p->ops)
                    if (o.tkey == key) adj_ok = adj_ok && o.adj_ok;
            for (const RegVariant* v : find_reg_variants(ref->prec, (int)ref->n, ref->ld_lf, ref->st_lf, ref->ldm,
                                                         ref->stm, ref->post, ref->fold ? ref->rv->fold : 0)) {
                if ((v->vec == 2 && !ref->pairable && !v->fold) || v->lds_bytes > kLdsBudget || (v->adj && !adj_ok))
                    continue;
                Op c = *ref;
                c.kind = OP_REG;
                c.rv = v;
                c.a.tw = TB.stage_twiddles(T.arith, Builder::parse_radices(v->radices));
                cands.emplace_back(reg_variant_name(v), c);
            }
            const uint64_t lstr = lds_line_stride(ref->n);
            const size_t es = elem_bytes(ref->prec == P_F32 ? P_F32 : P_F64);
            for (uint32_t lpw = 1; lpw <= 32 && !ref->fold; lpw *= 2) {
                if (2 * lstr * es * lpw > T.lds_budget || lpw > ref->a.lines) break;
                Op c = *ref;
                c.kind = OP_LDS;
                c.rv = nullptr;
                c.sc = make_sched(ref->n, lpw, ref->ld_lf, ref->st_lf);
                c.a.tw = TB.stage_twiddles(ref->prec == P_F32 ? P_F32 : P_F64, Builder::sched_radices(c.sc));
                cands.emplace_back("lds:lpw=" + std::to_string(lpw), c);
            }
            if (cands.size() < 2) continue;
            auto apply = [&](const Op& c) {
                for (auto& pv : ops)
                    for (Op& o : pv)
                        if (o.tkey == key) {
                            // keep each op's own buffers, lines and flags; take the kernel choice
                            o.kind = c.kind;
                            o.rv = c.rv;
                            o.sc = c.sc;
                            o.a.tw = c.a.tw;
                        }
            };
            hip_check(hipStreamSynchronize(s), "autotune tables");
            // three trials with the candidates interleaved; a candidate's time
            // is its best trial mean per execution of the sequence
            std::vector<float> tmin(cands.size(), 1e30f);
            for (int trial = 0; trial < 3; trial++) {
                for (size_t ci = 0; ci < cands.size(); ci++) {
                    apply(cands[ci].second);
                    for (int w = 0; w < 2; w++) run_all(w);
                    hip_check(hipEventRecord(e0, s), "autotune event");
                    for (int r = 0; r < reps; r++) run_all(r);
                    hip_check(hipEventRecord(e1, s), "autotune event");
                    hip_check(hipEventSynchronize(e1), "autotune event");
                    float ms = 0.0f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "autotune event");
                    tmin[ci] = std::min(tmin[ci], ms / (float)reps);
                }
            }
            size_t bi = 0;
            std::ostringstream line;
            line << key << ":";
            for (size_t ci = 0; ci < cands.size(); ci++) {
                line << " " << cands[ci].first << "=" << tmin[ci] * 1e3f << "us";
                if (tmin[ci] < tmin[bi]) bi = ci;
            }
            apply(cands[bi].second);
            tuning_set(key, cands[bi].first);
            tuned++;
            if (report) *report += line.str() + " -> " + cands[bi].first + "\n";
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    hip_check(hipStreamSynchronize(s), "autotune");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return tuned;
}

int autotune(const mifft_config& cfg, mifft_kind kind, int reps, hipStream_t s, std::string* report) {
    return autotune_sequence(&cfg, &kind, 1, reps, s, report);
}

}  // namespace mifft
