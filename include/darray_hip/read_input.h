// DArray's LIBSVM reader (read_input.h:8-38) over the library's host reader.
//
// Same signature and meaning: n points of d features into X (point j in
// column j, X[(index-1) + j*ldx]) and labels Y (1 -> +1, anything else ->
// -1).  X and Y are zeroed first; a feature index above d or a malformed
// token stops the program with a message (the reference writes past the
// column or ignores the token), and an unreadable file exits with
// EXIT_FAILURE as in the reference.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../darray_hip.h"

template <typename T>
void read_input_file(char* filename, int n, int d, T* X, int ldx, T* Y) {
    std::vector<double> x((size_t)(ldx > 0 ? ldx : 1) * (n > 0 ? n : 1)), y(n > 0 ? n : 1);
    int got = dh_read_libsvm(filename, n, d, x.data(), ldx, y.data());
    if (got < 0) {
        std::fprintf(stderr, "read_input_file: %s\n", dh_last_error_string());
        std::exit(EXIT_FAILURE);
    }
    for (int j = 0; j < n; ++j) {
        Y[j] = (T)y[j];
        for (int f = 0; f < d; ++f) X[f + (size_t)j * ldx] = (T)x[f + (size_t)j * ldx];
    }
}
