// Plain data types shared by host and device code.
#pragma once

namespace llamd {

struct cplx {
  double x, y;
};

}  // namespace llamd
