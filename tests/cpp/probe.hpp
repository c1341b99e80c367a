// A small check recorder for the C++ API test program (GoogleTest is not
// available offline).  Each suite is a function taking a Probe; checks record
// failures with the source location and carry on, `need` stops the suite.
// main() runs the registered suites, optionally only those that need no
// device (--cpu-only) or whose name contains a filter, and prints
// "<suites> suites, <n> failed checks".
#pragma once

#include <cmath>
#include <cstdio>
#include <exception>
#include <functional>
#include <source_location>
#include <string>
#include <utility>
#include <vector>

namespace probe {

struct Stop {};

class Probe {
public:
    // records a failed check unless ok
    bool expect(bool ok, const std::string& what, std::source_location at = std::source_location::current()) {
        if (!ok) record(what, at);
        return ok;
    }
    // as expect, and ends the suite on failure
    void need(bool ok, const std::string& what, std::source_location at = std::source_location::current()) {
        if (!expect(ok, what, at)) throw Stop{};
    }
    bool near(double a, double b, double tol, const std::string& what,
              std::source_location at = std::source_location::current()) {
        const bool ok = std::fabs(a - b) <= tol;
        if (!ok) record(what + " (" + std::to_string(a) + " vs " + std::to_string(b) + ")", at);
        return ok;
    }
    // fn must throw E (or a type derived from it)
    template <class E, class F>
    bool raises(F&& fn, const std::string& what, std::source_location at = std::source_location::current()) {
        try {
            fn();
        }
        catch (const E&) {
            return true;
        }
        catch (const std::exception& e) {
            record(what + ": threw another exception: " + e.what(), at);
            return false;
        }
        record(what + ": did not throw", at);
        return false;
    }
    template <class F>
    bool quiet(F&& fn, const std::string& what, std::source_location at = std::source_location::current()) {
        try {
            fn();
            return true;
        }
        catch (const std::exception& e) {
            record(what + ": threw " + e.what(), at);
            return false;
        }
    }
    int failed() const { return failed_; }

private:
    void record(const std::string& what, const std::source_location& at) {
        ++failed_;
        std::fprintf(stderr, "%s:%u: check failed: %s\n", at.file_name(), unsigned(at.line()), what.c_str());
    }
    int failed_ = 0;
};

struct Suite {
    const char* name;
    bool device;
    std::function<void(Probe&)> body;
};

inline int run(const std::vector<Suite>& suites, int argc, char** argv) {
    bool hostOnly = false;
    std::string filter;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--cpu-only")
            hostOnly = true;
        else
            filter = arg;
    }
    Probe p;
    int ran = 0;
    for (const auto& s : suites) {
        if ((hostOnly && s.device) || (!filter.empty() && std::string(s.name).find(filter) == std::string::npos))
            continue;
        const int before = p.failed();
        std::printf("suite %s ...\n", s.name);
        std::fflush(stdout);
        try {
            s.body(p);
        }
        catch (const Stop&) {
        }
        catch (const std::exception& e) {
            p.expect(false, std::string("uncaught exception in suite: ") + e.what());
        }
        std::printf("suite %s: %s\n", s.name, p.failed() == before ? "ok" : "FAILED");
        std::fflush(stdout);
        ++ran;
    }
    std::printf("%d suites, %d failed checks\n", ran, p.failed());
    return p.failed() ? 1 : 0;
}

}  // namespace probe
