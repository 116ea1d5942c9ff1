// A user program written against the drop-in C++ headers: the reference's
// wave_test (kernel.cu:45-64) -- AMR<wave_AMR_data> with 128 slots,
// init_temp, the root's 8 children, copy_bounds, then timestep() with a
// slice rasterised and saved as a PNG after every step.  Only the include
// lines and the geometry argument differ from the reference's program.
//
//   wave_plugin_demo <out_dir> <steps> <image_width> <image_height>
//
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
    if (argc != 5) {
        fprintf(stderr, "usage: %s out_dir steps width height\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int steps = atoi(argv[2]);
    const uint32_t W = (uint32_t)atoi(argv[3]), H = (uint32_t)atoi(argv[4]);
    const uint32_t slots = 128u;
    am_geometry g;
    if (am_geometry_init(&g, 16, 3, 6, 100.f, .4f)) return 3;
    try {
        AMR<wave_AMR_data> amr(slots, g);
        const size_t n = wave_AMR_data::cells(g) * slots;
        amr.sim_data.old_field.zero();
        amr.sim_data.new_field.zero();
        init_temp<wave_AMR_data>(amr.sim_data.old_field, amr.sim_data.new_field, amr);
        for (uint32_t c = 0; c < 8; c++) amr.add_node(0, c);
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
