// Large-block kernels (ny > 64): placeholder until the tiled condensation
// and cyclic reduction are in place.
#pragma once
