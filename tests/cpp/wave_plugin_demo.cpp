// A user program written against the drop-in C++ headers, shaped like the
// reference's wave_test (kernel.cu:45-64): an AMR<wave_AMR_data>, children
// added under the root, copy_bounds, then timestep() with a slice rasterised
// and saved as a PNG after every step.  Only the include lines and the
// geometry argument differ from a program written for the reference.
//
//   wave_plugin_demo <init.bin> <out_dir> <steps> <image_width> <image_height>
//
// init.bin holds old_field = new_field for all slots (max_slots * T^3 fp32).
// Writes <out_dir>/wave_<i>.png for i = 0..steps, <out_dir>/old.bin and
// new.bin, and prints "seed <s> used <n> timer <t>".
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include <adaptivemesh/AMR.hpp>
#include <adaptivemesh/SimpleWave.hpp>
#include <adaptivemesh/image.hpp>

using namespace am;

static bool dump(const char *path, const float *dev, size_t n)
{
    std::vector<float> h(n);
    if (am_memcpy_d2h(h.data(), dev, n * sizeof(float), nullptr)) return false;
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    const bool ok = fwrite(h.data(), sizeof(float), n, f) == n;
    return fclose(f) == 0 && ok;
}

int main(int argc, char **argv)
{
    if (argc != 6) {
        fprintf(stderr, "usage: %s init.bin out_dir steps width height\n", argv[0]);
        return 2;
    }
    const std::string out = argv[2];
    const int steps = atoi(argv[3]);
    const uint32_t W = (uint32_t)atoi(argv[4]), H = (uint32_t)atoi(argv[5]);
    const uint32_t slots = 40u;
    am_geometry g;
    if (am_geometry_init(&g, 16, 3, 6, 100.f, .4f)) return 3;
    try {
        AMR<wave_AMR_data> amr(slots, g);
        const size_t n = wave_AMR_data::cells(g) * slots;
        std::vector<float> init(n);
        FILE *f = fopen(argv[1], "rb");
        if (!f || fread(init.data(), sizeof(float), n, f) != n) {
            fprintf(stderr, "cannot read %s\n", argv[1]);
            return 4;
        }
        fclose(f);
        check(am_memcpy_h2d(amr.sim_data.old_field.gpu_buffer_ptr, init.data(), n * sizeof(float), nullptr), "init");
        check(am_memcpy_h2d(amr.sim_data.new_field.gpu_buffer_ptr, init.data(), n * sizeof(float), nullptr), "init");
        check(am_stream_sync(nullptr), "init");

        for (int c = 0; c < 8; c++) amr.add_node(0, c);  // the mesh of the tests' graded recipe
        for (int c = 0; c < 8; c++) amr.add_node(1, c);
        for (int c = 0; c < 8; c++) amr.add_node(16, c);
        amr.sim_data.copy_bounds();

        smart_gpu_cpu_buffer<rgba_f32> temp((size_t)W * H);
        smart_gpu_cpu_buffer<uint32_t> temp2((size_t)W * H);
        rasterise_to_image(amr.sim_data.new_field, temp, amr, W, H);
        save_image(temp2, temp, W, H, (out + "/wave_0.png").c_str());
        for (int i = 1; i <= steps; i++) {
            amr.sim_data.timestep();
            rasterise_to_image(amr.sim_data.new_field, temp, amr, W, H);
            save_image(temp2, temp, W, H, (out + "/wave_" + std::to_string(i) + ".png").c_str());
        }
        if (!dump((out + "/old.bin").c_str(), amr.sim_data.old_field.gpu_buffer_ptr, n) ||
            !dump((out + "/new.bin").c_str(), amr.sim_data.new_field.gpu_buffer_ptr, n))
            return 5;
        printf("seed %d used %u timer %u\n", amr.sim_data.rng.seed, amr.curr_used_slots(), amr.timer_helper);
        temp.destroy();
        temp2.destroy();
        amr.sim_data.destroy();
        amr.destroy();
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
