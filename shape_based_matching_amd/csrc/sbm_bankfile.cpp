// Packed binary bank file: a detector's whole template bank and its parameters in one file, so that a
// serving detector starts without parsing YAML.  It sits beside the reference's persistence
// (readClasses / writeClasses, line2Dup.h:183-185; Detector::read / write, line2Dup.cpp:1400-1415),
// which keeps its YAML files unchanged.  Paths ending in ".gz" are gzip-compressed; reading accepts
// plain or gzip files.
//
// Layout (little-endian):
//   "SBMBANK1"  u32 version (1)  u32 orientations  u32 levels  u32 T[levels]
//   f32 weak_threshold  f32 strong_threshold  u32 num_features  u32 classes
//   per class (std::map order): u32 name bytes, name, u32 templates,
//     per template, per level: i32 width, height, tl_x, tl_y, pyramid_level, u32 features,
//       per feature: i16 x, i16 y, u16 label, u16 0, f32 theta
// Feature angles (theta) are kept, so a loaded bank can still be rotated with addTemplate_rotate;
// the YAML files do not hold them.
#include <cstring>

#include "sbm_detector.h"

namespace sbm {

namespace {
constexpr char kMagic[8] = {'S', 'B', 'M', 'B', 'A', 'N', 'K', '1'};
constexpr uint32_t kVersion = 1;

struct Writer {
    std::string out;
    void raw(const void* p, size_t n) { out.append(static_cast<const char*>(p), n); }
    void u32(uint32_t v) { raw(&v, 4); }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
};

struct Reader {
    const std::string& s;
    size_t p = 0;
    explicit Reader(const std::string& str) : s(str) {}
    void raw(void* dst, size_t n) {
        if (n > s.size() - p) throw Error(kEIO, "bank file truncated");
        memcpy(dst, s.data() + p, n);
        p += n;
    }
    uint32_t u32() { uint32_t v; raw(&v, 4); return v; }
    int32_t i32() { int32_t v; raw(&v, 4); return v; }
    float f32() { float v; raw(&v, 4); return v; }
    size_t left() const { return s.size() - p; }
};

struct FeatureRecord {
    int16_t x, y;
    uint16_t label, zero;
    float theta;
};
static_assert(sizeof(FeatureRecord) == 12, "packed feature record");
}  // namespace

std::string Detector::write_bank_binary() const {
    Writer w;
    w.raw(kMagic, sizeof(kMagic));
    w.u32(kVersion);
    w.u32((uint32_t)orientation_bins);
    w.u32((uint32_t)pyramid_levels);
    for (int t : T_at_level) w.u32((uint32_t)t);
    w.f32(weak_threshold);
    w.f32(strong_threshold);
    w.u32((uint32_t)num_features);
    w.u32((uint32_t)class_templates.size());
    std::vector<FeatureRecord> recs;
    for (const auto& kv : class_templates) {
        w.u32((uint32_t)kv.first.size());
        w.raw(kv.first.data(), kv.first.size());
        w.u32((uint32_t)kv.second.size());
        for (const TemplatePyramid& tp : kv.second) {
            if ((int)tp.size() != pyramid_levels)
                throw Error(kEInval, "template pyramid of class '" + kv.first + "' has " + std::to_string(tp.size()) +
                                         " levels, the detector has " + std::to_string(pyramid_levels));
            for (const Template& te : tp) {
                w.i32(te.width);
                w.i32(te.height);
                w.i32(te.tl_x);
                w.i32(te.tl_y);
                w.i32(te.pyramid_level);
                w.u32((uint32_t)te.features.size());
                recs.resize(te.features.size());
                for (size_t i = 0; i < te.features.size(); ++i) {
                    const Feature& f = te.features[i];
                    if (f.x < -32768 || f.x > 32767 || f.y < -32768 || f.y > 32767)
                        throw Error(kEInval, "feature coordinate outside the int16 range");
                    if (f.label < 0 || f.label > 65535) throw Error(kEInval, "Invalid value of quantized parameter");
                    recs[i] = FeatureRecord{(int16_t)f.x, (int16_t)f.y, (uint16_t)f.label, 0, f.theta};
                }
                w.raw(recs.data(), recs.size() * sizeof(FeatureRecord));
            }
        }
    }
    return w.out;
}

// Everything is checked before the detector changes: a bad file leaves it as it was.
int Detector::read_bank_binary(const std::string& data) {
    Reader r(data);
    char magic[8];
    r.raw(magic, sizeof(magic));
    if (memcmp(magic, kMagic, sizeof(magic)) != 0) throw Error(kEIO, "not a bank file");
    if (r.u32() != kVersion) throw Error(kEIO, "unsupported bank file version");
    const uint32_t bins = r.u32(), levels = r.u32();
    if (levels > 16) throw Error(kEIO, "bank file: bad pyramid");
    std::vector<int> T(levels);
    for (auto& t : T) t = (int)r.u32();
    if ((int)bins != orientation_bins || (int)levels != pyramid_levels || T != T_at_level)
        throw Error(kEInval, "bank file made for another detector (orientations or pyramid differ)");
    const float weak = r.f32(), strong = r.f32();
    const int nf = (int)r.u32();
    const uint32_t nclasses = r.u32();
    std::vector<std::pair<std::string, std::vector<TemplatePyramid>>> loaded;
    std::vector<FeatureRecord> recs;
    for (uint32_t c = 0; c < nclasses; ++c) {
        const uint32_t len = r.u32();
        if (len > r.left()) throw Error(kEIO, "bank file truncated");
        std::string name(len, '\0');
        r.raw(&name[0], len);
        if (class_templates.count(name)) throw Error(kEDupClass, "class already loaded: " + name);
        for (const auto& l : loaded)
            if (l.first == name) throw Error(kEIO, "bank file repeats class " + name);
        const uint32_t ntempl = r.u32();
        if (levels == 0 || ntempl > r.left() / (24 * levels)) throw Error(kEIO, "bank file truncated");
        std::vector<TemplatePyramid> tps(ntempl, TemplatePyramid(levels));
        for (auto& tp : tps)
            for (Template& te : tp) {
                te.width = r.i32();
                te.height = r.i32();
                te.tl_x = r.i32();
                te.tl_y = r.i32();
                te.pyramid_level = r.i32();
                const uint32_t n = r.u32();
                if (n > r.left() / sizeof(FeatureRecord)) throw Error(kEIO, "bank file truncated");
                recs.resize(n);
                r.raw(recs.data(), n * sizeof(FeatureRecord));
                te.features.resize(n);
                for (uint32_t i = 0; i < n; ++i) {
                    if (recs[i].label >= bins) throw Error(kEIO, "bank file: label out of range");
                    te.features[i].x = recs[i].x;
                    te.features[i].y = recs[i].y;
                    te.features[i].label = recs[i].label;
                    te.features[i].theta = recs[i].theta;
                }
            }
        loaded.emplace_back(std::move(name), std::move(tps));
    }
    if (r.left() != 0) throw Error(kEIO, "bank file has trailing bytes");
    weak_threshold = weak;
    strong_threshold = strong;
    num_features = nf;
    for (auto& l : loaded) class_templates.emplace(std::move(l.first), std::move(l.second));
    bank_dirty_ = true;
    return (int)nclasses;
}

}  // namespace sbm
