// Definitions of the opaque C-ABI handles declared in gb.h.
#pragma once

#include <memory>

#include "device_graph.hpp"
#include "host_graph.hpp"

struct gb_graph {
    gg::HostGraph g;
};

struct gb_dgraph {
    std::unique_ptr<gg::DeviceGraph> d;
};
