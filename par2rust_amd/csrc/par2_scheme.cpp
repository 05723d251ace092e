// par2_scheme.cpp -- volume split and volume file names of run_create.
// Host only: no device code and no HIP calls, so it also builds with a
// plain C++ compiler (the sanitizer harness does).
#include <algorithm>
#include <string>
#include <vector>

#include "par2_create.hpp"

namespace par2 {

static int scheme_err(Error &err, const std::string &msg) {
    err.code = P2G_ERR_INVALID_VOLUME_SCHEME;
    err.message = msg;
    return err.code;
}

// cap_volume_sizes, src/creator.rs:414-431: every entry above `cap` becomes
// cap-sized chunks and the remainder, in place.
static std::vector<uint32_t> cap_sizes(const std::vector<uint32_t> &in, uint32_t cap) {
    std::vector<uint32_t> out;
    for (uint32_t n : in) {
        uint32_t rem = n;
        while (rem > cap) {
            out.push_back(cap);
            rem -= cap;
        }
        if (rem > 0) out.push_back(rem);
    }
    return out;
}

// resolve_volume_sizes, src/creator.rs:325-362.  Limited wrappers nest
// (`inner: Box<VolumeScheme>`): s.caps holds their caps from the innermost
// outward.  The reference checks a wrapper's cap before resolving its inner
// scheme, so a zero cap is reported from the outermost wrapper inward, ahead
// of any error of the base scheme; the caps then apply from the innermost.
int resolve_volume_sizes(const VolumeScheme &s, uint32_t total, std::vector<uint32_t> &out,
                         Error &err) {
    out.clear();
    for (size_t k = s.caps.size(); k-- > 0;)
        if (s.caps[k] == 0) return scheme_err(err, "Limited scheme requires max_blocks_per_volume > 0");
    std::vector<uint32_t> inner;
    switch (s.kind) {
    case P2G_SCHEME_SINGLE:
        inner.push_back(total);
        break;
    case P2G_SCHEME_EXPONENTIAL: {  // exponential_split, src/creator.rs:367-385
        uint32_t remaining = total, next = 1;
        while (remaining > 0) {
            const uint32_t take = std::min(next, remaining);
            inner.push_back(take);
            remaining -= take;
            if (inner.size() >= 2) next = next > 0x7FFFFFFFu ? 0xFFFFFFFFu : next * 2;
        }
        break;
    }
    case P2G_SCHEME_UNIFORM: {  // uniform_split, src/creator.rs:390-409
        const uint32_t count = s.uniform_count;
        if (count == 0) return scheme_err(err, "Uniform scheme requires count > 0");
        if (count > total)
            return scheme_err(err, "Uniform scheme count (" + std::to_string(count) +
                                       ") exceeds recovery block total (" + std::to_string(total) + ")");
        const uint32_t base = total / count, rem = total - base * count;
        for (uint32_t i = 0; i + 1 < count; i++) inner.push_back(base);
        inner.push_back(base + rem);
        break;
    }
    case P2G_SCHEME_EXPLICIT: {  // src/creator.rs:346-360
        if (s.explicit_sizes.empty()) return scheme_err(err, "explicit scheme requires at least one volume");
        uint64_t sum = 0;
        for (uint32_t v : s.explicit_sizes) {
            if (v == 0) return scheme_err(err, "explicit scheme volumes must each have >0 recovery blocks");
            sum += v;
        }
        if (sum != total)
            return scheme_err(err, "explicit volume sizes sum to " + std::to_string(sum) + ", expected " +
                                       std::to_string(total));
        inner = s.explicit_sizes;
        break;
    }
    default:
        return scheme_err(err, "unknown volume scheme");
    }
    for (uint32_t cap : s.caps) inner = cap_sizes(inner, cap);
    out = inner;
    return P2G_OK;
}

// derive_volume_filename, src/creator.rs:1046-1057, with Rust's
// Path::file_stem / extension split (a leading dot does not start an
// extension).
std::string volume_name(const std::string &index, uint32_t first, uint32_t count) {
    const size_t slash = index.rfind('/');
    const size_t fstart = slash == std::string::npos ? 0 : slash + 1;
    const std::string dir = index.substr(0, fstart);
    const std::string fname = index.substr(fstart);
    const size_t dot = fname.rfind('.');
    const bool has_ext = dot != std::string::npos && dot != 0;
    const std::string stem = has_ext ? fname.substr(0, dot) : fname;
    const std::string ext = has_ext ? fname.substr(dot + 1) : "par2";
    return dir + stem + ".vol" + std::to_string(first) + "+" + std::to_string(count) + "." + ext;
}

}  // namespace par2
