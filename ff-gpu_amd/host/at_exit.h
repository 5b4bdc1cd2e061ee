// Scope guard: runs a release callback when its scope ends, so that error
// returns free device resources as the normal path does.
#pragma once

#include <functional>
#include <utility>

namespace ffg {

class AtExit {
public:
    explicit AtExit(std::function<void()> f) : f_(std::move(f)) {}
    ~AtExit() { f_(); }
    AtExit(const AtExit &) = delete;
    AtExit &operator=(const AtExit &) = delete;

private:
    std::function<void()> f_;
};

}  // namespace ffg
