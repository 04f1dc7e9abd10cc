// Opaque C handles. A handle either owns its object or borrows one owned by
// a beam (fa_beam_engine / fa_beam_rfi); borrowed handles are not destroyed.
#pragma once

#include <memory>

#include "engine.hpp"
#include "rfi.hpp"

struct fa_engine {
    std::unique_ptr<fa::Engine> owned;
    fa::Engine* eng = nullptr;
};

struct fa_rfi {
    std::unique_ptr<fa::DeviceRfi> owned;
    fa::DeviceRfi* rfi = nullptr;
};
