// dpd_main.cpp -- command-line driver: `dpd_amd <runId> [runId ...]` runs dmpci.<runId> from
// the current directory on the GPU, as the reference's `dpd <runId>` does (src/dmpc.cpp:64-226).
#include <cstdio>
#include <cstdlib>

#include "dpd_host.h"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <runId> [runId ...]\n", argv[0]);
        return 1;
    }
    const char* dev = std::getenv("DPD_DEVICE");
    const int device = dev ? std::atoi(dev) : 0;
    int status = 0;
    for (int i = 1; i < argc; ++i) {
        dpd_run_stats st;
        char err[512] = {0};
        const int rc = dpd_run(".", argv[i], device, &st, err, sizeof(err));
        if (rc != DPD_OK) {
            std::fprintf(stderr, "run %s failed (%d): %s\n", argv[i], rc, err);
            status = 2;
            continue;
        }
        std::printf("run %s: %lld steps, %lld beads, %.3f s, %.4g bead-steps/s, kT %.5f, P %.5f\n", argv[i],
                    (long long)st.steps, (long long)st.nbeads, st.seconds, st.bead_steps_per_s, st.last_kT,
                    st.last_pressure);
    }
    return status;
}
