// Opaque handles of the host C wrapper (include/spec_spars_host.h).
#pragma once
#include "../../include/spec_spars.hpp"

struct ssh_sparsifier {
    spec_spars::Sparsifier sp;
};
struct ssh_matrix {
    spec_spars::MtxMatrix m;
};
struct ssh_stream {
    spec_spars::InputStream is;
};
