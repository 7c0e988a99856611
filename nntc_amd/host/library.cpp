// nntc_encode_library: a library of materials encoded side by side on one GPU.
//
// The reference encodes one material per process (src/main.cpp:1753), and its CUDA arm overlaps work only inside a
// material, across the planes of the mip chain (src/device.cuh:269-301). A library of small and mid-sized materials
// leaves most of an MI355X idle that way: a 256 x 256 material's launches cover a few hundred workgroups against the
// chip's 256 CUs, and the host's decisions between launches (the CG probes, the decoder's ridge ladder, the
// objective's readbacks) leave the queue empty while they are taken. Here `workers` host threads each own one encoder
// at a time - one context with its own non-blocking stream - and take the next material from a queue ordered largest
// first, so the launches of several materials are in flight at once and one material's host decisions overlap the
// others' kernels. Each material runs the launches an encode of it alone runs, in the same order on its own stream, so
// its outputs are that encode's bits.

#include "nntc_host.h"

#include <algorithm>
#include <atomic>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace
{

std::mutex g_first_lock;

}   // namespace

extern "C" int nntc_encode_library(const nntc_material* materials, int count, const nntc_encode_options* o,
                                   int workers, nntc_encode_report* reports, int32_t* status)
{
    if (count <= 0)
        return 0;
    if (workers < 1)
        workers = 1;
    workers = std::min(workers, count);
    // Largest first (texels times textures), so the long materials start early and the short ones fill the end.
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double wa = (double)materials[a].width * materials[a].height * materials[a].textures;
        const double wb = (double)materials[b].width * materials[b].height * materials[b].textures;
        return wa > wb;
    });
    std::atomic<int> next { 0 };
    std::atomic<int> failed { 0 };
    std::string first_error;
    auto work = [&] {
        for (;;)
        {
            const int slot = next.fetch_add(1);
            if (slot >= count)
                return;
            const int i = order[(size_t)slot];
            const nntc_material& m = materials[i];
            nntc_encode_options opts = *o;
            opts.c0 = m.c0;
            opts.c1 = m.c1;
            nntc_encoder* e = m.on_device
                                  ? nntc_encoder_create_device(&opts, m.textures, m.width, m.height, m.rgb)
                                  : nntc_encoder_create(&opts, m.textures, m.width, m.height, m.rgb, nullptr, nullptr);
            int st = e ? nntc_encoder_run(e) : 1;
            if (st == 0 && m.prefix)
                st = nntc_encoder_write(e, m.prefix);
            if (st == 0 && reports)
                st = nntc_encoder_report(e, &reports[i]);
            if (st != 0)
            {
                failed++;
                std::lock_guard<std::mutex> g(g_first_lock);
                if (first_error.empty())
                    first_error = std::string("material ") + std::to_string(i) + ": " + nntc_encode_last_error();
            }
            if (status)
                status[i] = st;
            nntc_encoder_destroy(e);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < workers; w++)
        pool.emplace_back(work);
    work();
    for (std::thread& t : pool)
        t.join();
    if (!first_error.empty())
        nntc_host::g_error = first_error;
    return failed.load();
}
