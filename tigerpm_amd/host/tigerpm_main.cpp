// tigerpm_main.cpp -- command-line driver with the reference's --test=
// scenarios (main.cpp:9-18 and test.cpp:25-51), on one GPU.
#include <cstdio>
#include <exception>

#include "tigerpm_ref.hpp"

int main(int argc, char* argv[]) {
    try {
        if (!process_options(argc, argv)) return 1;
        run_test(get_options().test);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "tigerpm: %s\n", e.what());
        return 1;
    }
    return 0;
}
