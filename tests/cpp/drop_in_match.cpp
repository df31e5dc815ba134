// A C++ caller written against the reference's API, compiled against include/line2Dup.h.
// Usage: drop_in_match <bank.yaml[.gz]> <image.raw> <rows> <cols> <channels> <threshold>
// Prints one line per match: x y similarity(bits as hex) class_id template_id
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <vector>

#include "line2Dup.h"

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s bank image.raw rows cols channels threshold\n", argv[0]);
        return 2;
    }
    const int rows = std::atoi(argv[3]), cols = std::atoi(argv[4]), ch = std::atoi(argv[5]);
    const float thr = (float)std::atof(argv[6]);
    std::ifstream f(argv[2], std::ios::binary);
    std::vector<uint8_t> img((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if ((int)img.size() != rows * cols * ch) {
        std::fprintf(stderr, "image size mismatch\n");
        return 2;
    }
    try {
        line2Dup::Detector detector(63, {4, 8});
        const std::string cid = detector.readClass(argv[1]);
        std::vector<std::string> ids{cid};
        auto matches = detector.match(line2Dup::ImageView(img.data(), rows, cols, ch), thr, ids);
        for (const auto& m : matches) {
            uint32_t bits;
            std::memcpy(&bits, &m.similarity, 4);
            std::printf("%d %d %08x %s %d\n", m.x, m.y, bits, m.class_id.c_str(), m.template_id);
        }
        auto t = detector.getTemplates(cid, 0);
        std::fprintf(stderr, "templates=%d levels=%d T0=%d width0=%d matches=%zu\n", detector.numTemplates(),
                     detector.pyramidLevels(), detector.getT(0), t[0].width, matches.size());
    } catch (const line2Dup::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
