// trace.h -- roctx ranges named after the reference's tracing spans (trace.hip).
#pragma once

namespace wl {

void span_push(const char *name);
void span_pop();

// a range over a scope
struct Span {
    explicit Span(const char *name) { span_push(name); }
    ~Span() { span_pop(); }
    Span(const Span &) = delete;
    Span &operator=(const Span &) = delete;
};
// consecutive phases of one scope: next() ends the current range and opens another
struct PhaseSpan {
    explicit PhaseSpan(const char *name) { span_push(name); }
    void next(const char *name) {
        span_pop();
        span_push(name);
    }
    ~PhaseSpan() { span_pop(); }
    PhaseSpan(const PhaseSpan &) = delete;
    PhaseSpan &operator=(const PhaseSpan &) = delete;
};

}  // namespace wl
