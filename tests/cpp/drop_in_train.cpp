// A C++ training loop written against the reference's API (the shape of test.cpp:385-410), compiled
// against include/line2Dup.h.
// Usage: drop_in_train <image.raw> <rows> <cols> <channels> <num_features> <angle_step> <out_prefix>
// Writes <out_prefix>test_templ.yaml and <out_prefix>test_info.yaml; prints the template ids.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <vector>

#include "line2Dup.h"

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s image.raw rows cols channels num_features angle_step out_prefix\n", argv[0]);
        return 2;
    }
    const int rows = std::atoi(argv[2]), cols = std::atoi(argv[3]), ch = std::atoi(argv[4]);
    const int nf = std::atoi(argv[5]);
    const float step = (float)std::atof(argv[6]);
    const std::string prefix = argv[7];
    std::ifstream f(argv[1], std::ios::binary);
    std::vector<uint8_t> img((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if ((int)img.size() != rows * cols * ch) {
        std::fprintf(stderr, "image size mismatch\n");
        return 2;
    }
    try {
        line2Dup::Detector detector(nf, {4, 8});
        line2Dup::ImageView src(img.data(), rows, cols, ch);
        std::vector<uint8_t> m((size_t)rows * cols, 255);
        line2Dup::ImageView mask(m.data(), rows, cols, 1);

        shape_based_matching::shapeInfo_producer shapes(src, mask);
        shapes.angle_range = {0, 360};
        shapes.angle_step = step;
        shapes.produce_infos();
        std::vector<shape_based_matching::shapeInfo_producer::Info> infos_have_templ;
        std::string class_id = "test";
        for (auto& info : shapes.infos) {
            int templ_id = detector.addTemplate(shapes.src_of(info), class_id, shapes.mask_of(info));
            std::printf("%d\n", templ_id);
            if (templ_id != -1) infos_have_templ.push_back(info);
        }
        detector.writeClasses(prefix + "%s_templ.yaml");
        shapes.save_infos(infos_have_templ, prefix + "test_info.yaml");
    } catch (const line2Dup::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
