// verbosity.h -- verbosity levels and settings of the solvers, as in the
// reference (verbosity.h:9-24, 27-70).  The solver library prints the
// reference's lines itself; these helpers build the settings passed in.
#ifndef GLX_COMPAT_VERBOSITY
#define GLX_COMPAT_VERBOSITY

#include <iostream>
#include <string>

enum inversion_verbose_level {
  VERB_PASS_THROUGH = -1,    // preconditioner: use the outer setting
  VERB_NONE = 0,             // print nothing
  VERB_SUMMARY = 1,          // one line at the end
  VERB_RESTART_DETAIL = 2,   // plus one line per restart
  VERB_DETAIL = 3            // plus one line per iteration
};

struct inversion_verbose_struct {
  inversion_verbose_level verbosity;
  std::string verb_prefix;
  inversion_verbose_level precond_verbosity;
  std::string precond_verb_prefix;
};

// Settings for the inner solve of a restarted solver: SUMMARY and
// RESTART_DETAIL are silenced there, DETAIL passes through.
inline void shuffle_verbosity_restart(inversion_verbose_struct* inner, inversion_verbose_struct* outer) {
  if (outer) {
    const bool quiet = outer->verbosity == VERB_RESTART_DETAIL || outer->verbosity == VERB_SUMMARY;
    inner->verbosity = quiet ? VERB_NONE : outer->verbosity;
    inner->verb_prefix = outer->verb_prefix;
    inner->precond_verbosity = outer->precond_verbosity;
    inner->precond_verb_prefix = outer->precond_verb_prefix;
  } else {
    inner->verbosity = VERB_NONE;
    inner->verb_prefix = "";
    inner->precond_verbosity = VERB_NONE;
    inner->precond_verb_prefix = "";
  }
}

// Settings handed to a preconditioner's inner solve.
inline void shuffle_verbosity_precond(inversion_verbose_struct* inner, inversion_verbose_struct* outer) {
  if (outer) {
    inner->verbosity = outer->precond_verbosity == VERB_PASS_THROUGH ? outer->verbosity : outer->precond_verbosity;
    inner->verb_prefix = outer->precond_verb_prefix;
  } else {
    inner->verbosity = VERB_NONE;
    inner->verb_prefix = "";
  }
  inner->precond_verbosity = VERB_NONE;
  inner->precond_verb_prefix = "";
}

#endif
