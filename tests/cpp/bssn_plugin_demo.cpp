// A user program written against the drop-in C++ headers: the reference's
// BSSN_test (kernel.cu:66-111) -- AMR<BSSN_AMR_data> with 128 slots, the
// root's 8 children refined towards the centre 5 more times, init_temp of
// the six metric components, two timesteps (the second timed), and the six
// Christoffel word arrays saved as images.  Only the include lines, the
// geometry argument and size_domain (a runtime value here) differ from the
// reference's program.
//
//   bssn_plugin_demo <out_dir>
//
// Also writes the state the test compares: <out_dir>/new.bin (the six
// components of cyij_new, used slots), cG.bin (the six Christoffel word
// arrays), cR.bin (the six metric traces), and prints
// "seed <s> used <n> timer <t>" as its last line.
#include <stdio.h>

#include <chrono>
#include <iostream>
#include <string>
#include <vector>

#include <adaptivemesh/AMR.hpp>
#include <adaptivemesh/BSSN.hpp>
#include <adaptivemesh/SimpleWave.hpp>
#include <adaptivemesh/image.hpp>

using namespace am;

template <class B>
static bool dump(FILE *f, B &buf, size_t n)
{
    std::vector<uint32_t> h(n);
    if (am_memcpy_d2h(h.data(), buf.gpu_buffer_ptr, n * 4, nullptr)) return false;
    return fwrite(h.data(), 4, n, f) == n;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s out_dir\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    am_geometry g;
    if (am_geometry_init(&g, 16, 3, 6, 100.f, .4f)) return 3;
    const uint32_t size_domain = (uint32_t)g.size_domain;
    try {
        AMR<BSSN_AMR_data> amr(128u, g);
        int children[8];

        for (uint32_t c = 0; c < 8; c++) {
            children[c] = (int)amr.read_first_free_slot();
            amr.add_node(0, (int)c);
        }

        for (uint32_t d = 1; d < 6; d++)
            for (uint32_t c = 0; c < 8; c++) {
                const uint32_t new_idx = amr.read_first_free_slot();
                amr.add_node(children[c], (int)((~c) & 7u));
                children[c] = (int)new_idx;
            }

        round_robin_threads threads;
        fast_prng rng;
        std::cout << amr.to_string_debug() + "\n\n";

        // dumb init for testing
        init_temp(amr.sim_data.cyij_old.xx, amr.sim_data.cyij_new.xx, amr, 1.f);
        init_temp(amr.sim_data.cyij_old.yy, amr.sim_data.cyij_new.yy, amr, 1.f);
        init_temp(amr.sim_data.cyij_old.zz, amr.sim_data.cyij_new.zz, amr, 1.f);
        init_temp(amr.sim_data.cyij_old.xy, amr.sim_data.cyij_new.xy, amr, 0.f);
        init_temp(amr.sim_data.cyij_old.xz, amr.sim_data.cyij_new.xz, amr, 0.f);
        init_temp(amr.sim_data.cyij_old.yz, amr.sim_data.cyij_new.yz, amr, 0.f);

        amr.sim_data.timestep();
        std::chrono::steady_clock clock;
        cuda_sync();
        long long now = clock.now().time_since_epoch().count();
        amr.sim_data.timestep();
        cuda_sync();
        long long now2 = clock.now().time_since_epoch().count();
        std::cout << (now2 - now) * 1E-9 << "\n";

        smart_gpu_cpu_buffer<uint32_t> temp(amr.sim_data.cGijk.xx.dedicated_len);
        const char *names[6] = {"xx", "xy", "xz", "yy", "yz", "zz"};
        for (int c = 0; c < 6; c++)
            save_image(temp, amr.sim_data.cGijk[c], size_domain * size_domain,
                       (uint32_t)(temp.dedicated_len / (size_domain * size_domain)),
                       (out + "/test_" + names[c] + ".png").c_str());

        const size_t used = amr.curr_used_slots();
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain), S = (size_t)g.size_domain;
        FILE *fn = fopen((out + "/new.bin").c_str(), "wb"), *fg = fopen((out + "/cG.bin").c_str(), "wb"),
             *fr = fopen((out + "/cR.bin").c_str(), "wb");
        if (!fn || !fg || !fr) return 4;
        for (int c = 0; c < 6; c++)
            if (!dump(fn, amr.sim_data.cyij_new[c], used * T * T * T) || !dump(fg, amr.sim_data.cGijk[c], used * S * S * S) ||
                !dump(fr, amr.sim_data.cRij[c], used * S * S * S))
                return 5;
        fclose(fn);
        fclose(fg);
        fclose(fr);
        printf("seed %d used %zu timer %u\n", amr.sim_data.rng.seed, used, amr.timer_helper);
        temp.destroy();
        amr.sim_data.destroy();
        amr.destroy();
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
