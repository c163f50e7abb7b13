#include "schedule.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>

namespace shgpu {

namespace {

shgpu_schedule_class cls(uint32_t variant, uint64_t max_keys, uint32_t waves) {
    shgpu_schedule_class c{};
    c.variant = variant;
    c.waves_per_cu = waves;
    c.sh1_seeds_per_lane = 4;
    c.split_seeds_per_lane = 2;
    c.max_keys = max_keys;
    return c;
}

bool less_class(const shgpu_schedule_class &a, const shgpu_schedule_class &b) {
    return a.variant != b.variant ? a.variant < b.variant : a.max_keys < b.max_keys;
}

}  // namespace

Schedule Schedule::builtin() {
    // Kernel ms of ShockHash2 leaf 64 (bucket 2000) at 4 / 8 / 12 / 16 / 24 waves per CU: 200k keys 2.5 / 2.8 /
    // 3.7 / 4.7 / 6.3; 1M 6.6 / 5.1 / 5.3 / 5.8 / 7.6; 2M 12.6 / 8.7 / 7.9 / 8.0 / 10.2; 10M - / 38.2 / 33.0 / 30.9
    // / 29.7. ShockHash leaf 30 has the same optimum at every size measured; SIMDShockHash leaves carry ~1000x the
    // work per key and keep 24 waves.
    Schedule s;
    for (uint32_t v : {(uint32_t)SHGPU_SHOCKHASH, (uint32_t)SHGPU_SHOCKHASH2}) {
        s.classes.push_back(cls(v, 350000, 4));
        s.classes.push_back(cls(v, 1500000, 8));
        s.classes.push_back(cls(v, 3000000, 12));
        s.classes.push_back(cls(v, 5000000, 16));
        s.classes.push_back(cls(v, UINT64_MAX, 24));
    }
    s.classes.push_back(cls(SHGPU_SIMD_SHOCKHASH, UINT64_MAX, 24));
    std::sort(s.classes.begin(), s.classes.end(), less_class);
    return s;
}

const shgpu_schedule_class &Schedule::lookup(uint32_t variant, uint64_t n) const {
    const shgpu_schedule_class *last = nullptr;
    for (const auto &c : classes) {
        if (c.variant != variant) continue;
        if (n <= c.max_keys) return c;
        last = &c;
    }
    return *last;  // every variant has a class (builtin, replace)
}

void Schedule::put(const shgpu_schedule_class &c) {
    for (auto &x : classes)
        if (x.variant == c.variant && x.max_keys == c.max_keys) {
            x = c;
            return;
        }
    classes.push_back(c);
    std::sort(classes.begin(), classes.end(), less_class);
}

int check_class(const shgpu_schedule_class &c, std::string &err) {
    auto seeds_ok = [](uint32_t s) { return s == 1 || s == 2 || s == 4; };
    if (c.variant > 2) { err = "schedule class: unknown variant " + std::to_string(c.variant); return SHGPU_ERR_ARG; }
    if (c.waves_per_cu < 1 || c.waves_per_cu > 32) { err = "schedule class: waves_per_cu must be 1..32"; return SHGPU_ERR_ARG; }
    if (!seeds_ok(c.sh1_seeds_per_lane) || !seeds_ok(c.split_seeds_per_lane)) {
        err = "schedule class: seeds per lane must be 1, 2 or 4";
        return SHGPU_ERR_ARG;
    }
    if (c.max_keys == 0) { err = "schedule class: max_keys must be positive"; return SHGPU_ERR_ARG; }
    return SHGPU_OK;
}

int Schedule::replace(const std::vector<shgpu_schedule_class> &cls_in, std::string &err) {
    bool named[3] = {false, false, false};
    for (const auto &c : cls_in) {
        int rc = check_class(c, err);
        if (rc) return rc;
        named[c.variant] = true;
    }
    std::vector<shgpu_schedule_class> next;
    for (const auto &c : classes)
        if (!named[c.variant]) next.push_back(c);
    next.insert(next.end(), cls_in.begin(), cls_in.end());
    std::sort(next.begin(), next.end(), less_class);
    for (size_t i = 1; i < next.size(); i++)
        if (next[i].variant == next[i - 1].variant && next[i].max_keys == next[i - 1].max_keys) {
            err = "schedule: two classes of variant " + std::to_string(next[i].variant) + " with max_keys " +
                  std::to_string(next[i].max_keys);
            return SHGPU_ERR_ARG;
        }
    classes = std::move(next);
    return SHGPU_OK;
}

std::string Schedule::to_json(const std::string &arch) const {
    std::string s = "{\n  \"arch\": \"" + arch + "\",\n  \"classes\": [\n";
    char buf[512];
    for (size_t i = 0; i < classes.size(); i++) {
        const auto &c = classes[i];
        std::snprintf(buf, sizeof(buf),
                      "    {\"variant\": %u, \"max_keys\": %llu, \"waves_per_cu\": %u, \"sh1_seeds_per_lane\": %u, "
                      "\"split_seeds_per_lane\": %u, \"tuned_ms\": %.17g, \"sample_keys\": %llu}%s\n",
                      c.variant, (unsigned long long)c.max_keys, c.waves_per_cu, c.sh1_seeds_per_lane,
                      c.split_seeds_per_lane, c.tuned_ms, (unsigned long long)c.sample_keys,
                      i + 1 < classes.size() ? "," : "");
        s += buf;
    }
    s += "  ]\n}\n";
    return s;
}

namespace {

// Minimal reader for the schedule's JSON: objects, arrays, strings without escapes, numbers.
struct Reader {
    const std::string &t;
    size_t i = 0;
    std::string err;
    explicit Reader(const std::string &text) : t(text) {}
    void ws() {
        while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
    }
    bool eat(char c) {
        ws();
        if (i < t.size() && t[i] == c) { i++; return true; }
        return false;
    }
    bool string(std::string &out) {
        ws();
        if (i >= t.size() || t[i] != '"') { err = "expected a string"; return false; }
        const size_t j = t.find('"', i + 1);
        if (j == std::string::npos) { err = "unterminated string"; return false; }
        out = t.substr(i + 1, j - i - 1);
        i = j + 1;
        return true;
    }
    bool number(std::string &out) {
        ws();
        const size_t j = i;
        while (i < t.size() && (std::isdigit((unsigned char)t[i]) || t[i] == '-' || t[i] == '+' || t[i] == '.' ||
                                t[i] == 'e' || t[i] == 'E'))
            i++;
        if (i == j) { err = "expected a number"; return false; }
        out = t.substr(j, i - j);
        return true;
    }
    // Skips any value (used for keys the reader does not know).
    bool skip() {
        ws();
        if (i >= t.size()) { err = "unexpected end"; return false; }
        std::string tmp;
        if (t[i] == '"') return string(tmp);
        if (t[i] == '{' || t[i] == '[') {
            const char open = t[i], close = open == '{' ? '}' : ']';
            i++;
            if (eat(close)) return true;
            do {
                if (open == '{') {
                    if (!string(tmp) || !eat(':')) { if (err.empty()) err = "expected ':'"; return false; }
                }
                if (!skip()) return false;
            } while (eat(','));
            if (!eat(close)) { err = "expected a closing bracket"; return false; }
            return true;
        }
        if (t.compare(i, 4, "true") == 0 || t.compare(i, 4, "null") == 0) { i += 4; return true; }
        if (t.compare(i, 5, "false") == 0) { i += 5; return true; }
        return number(tmp);
    }
};

}  // namespace

int parse_schedule_json(const std::string &text, std::vector<shgpu_schedule_class> &out, std::string &err) {
    out.clear();
    Reader r(text);
    auto bad = [&](const std::string &what) {
        err = "schedule JSON: " + what + " at offset " + std::to_string(r.i);
        return SHGPU_ERR_ARG;
    };
    if (!r.eat('{')) return bad("expected '{'");
    bool have_classes = false;
    if (!r.eat('}')) {
        do {
            std::string key;
            if (!r.string(key) || !r.eat(':')) return bad(r.err.empty() ? "expected ':'" : r.err);
            if (key != "classes") {
                if (!r.skip()) return bad(r.err);
                continue;
            }
            have_classes = true;
            if (!r.eat('[')) return bad("expected '['");
            if (r.eat(']')) continue;
            do {
                if (!r.eat('{')) return bad("expected '{'");
                shgpu_schedule_class c{};
                unsigned seen = 0;
                if (!r.eat('}')) {
                    do {
                        std::string k, num;
                        if (!r.string(k) || !r.eat(':')) return bad(r.err.empty() ? "expected ':'" : r.err);
                        static const char *names[] = {"variant", "max_keys", "waves_per_cu", "sh1_seeds_per_lane",
                                                      "split_seeds_per_lane", "tuned_ms", "sample_keys"};
                        int which = -1;
                        for (int q = 0; q < 7; q++)
                            if (k == names[q]) which = q;
                        if (which < 0) {
                            if (!r.skip()) return bad(r.err);
                            continue;
                        }
                        if (!r.number(num)) return bad(r.err);
                        char *end = nullptr;
                        if (which == 5) {
                            c.tuned_ms = std::strtod(num.c_str(), &end);
                        } else {
                            if (num.find_first_of("-.eE") != std::string::npos) return bad("'" + k + "' must be a non-negative integer");
                            const unsigned long long v = std::strtoull(num.c_str(), &end, 10);
                            if (which != 1 && which != 6 && v > 0xFFFFFFFFull) return bad("'" + k + "' out of range");
                            switch (which) {
                                case 0: c.variant = (uint32_t)v; break;
                                case 1: c.max_keys = v; break;
                                case 2: c.waves_per_cu = (uint32_t)v; break;
                                case 3: c.sh1_seeds_per_lane = (uint32_t)v; break;
                                case 4: c.split_seeds_per_lane = (uint32_t)v; break;
                                default: c.sample_keys = v; break;
                            }
                        }
                        if (!end || *end) return bad("bad number '" + num + "'");
                        seen |= 1u << which;
                    } while (r.eat(','));
                    if (!r.eat('}')) return bad("expected '}'");
                }
                if ((seen & 0x1F) != 0x1F) return bad("a class lacks variant, max_keys, waves_per_cu or a seeds-per-lane value");
                int rc = check_class(c, err);
                if (rc) return rc;
                out.push_back(c);
            } while (r.eat(','));
            if (!r.eat(']')) return bad("expected ']'");
        } while (r.eat(','));
        if (!r.eat('}')) return bad("expected '}'");
    }
    r.ws();
    if (r.i != text.size()) return bad("trailing characters");
    if (!have_classes) return bad("no \"classes\"");
    return SHGPU_OK;
}

}  // namespace shgpu
