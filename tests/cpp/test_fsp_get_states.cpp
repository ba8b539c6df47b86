// Mirrors the reference's tests/test_fsp_get_states.cu: enumerate the states
// of the box (1023, 1023, 3) on the GPU and check the order exhaustively.
#include <cassert>
#include <cstdio>
#include <vector>

#include "cme_util.h"

int main()
{
    int d = 3;
    int n_bounds[3] = {1023, 1023, 3};
    int n_states = cuFSP::rect_fsp_num_states(d, n_bounds);
    assert(n_states == 4194304);
    int *d_states = nullptr;
    cuFSP::throw_on(fsp_malloc((void **)&d_states, sizeof(int) * (size_t)d * n_states), "malloc");
    cuFSP::fsp_get_states(d_states, d, n_states, n_bounds);
    std::vector<int> states((size_t)d * n_states);
    cuFSP::throw_on(fsp_memcpy_d2h(states.data(), d_states, states.size() * sizeof(int)), "d2h");
    bool success = true;
    int i = 0;
    for (int x2 = 0; x2 <= n_bounds[2]; ++x2)
        for (int x1 = 0; x1 <= n_bounds[1]; ++x1)
            for (int x0 = 0; x0 <= n_bounds[0]; ++x0) {
                success = success && x0 == states[(size_t)i * d] && x1 == states[(size_t)i * d + 1] &&
                          x2 == states[(size_t)i * d + 2];
                int x[3];
                cuFSP::indx2state(i, x, d, n_bounds);
                success = success && cuFSP::state2indx(x, d, n_bounds) == i;
                ++i;
            }
    assert(success);
    fsp_free(d_states);
    std::printf("Test FSP state generation successful.\n");
    return 0;
}
