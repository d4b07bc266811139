// Declarations shared between the library's translation units (not part of
// the public C ABI).
#pragma once

#include "qmbed.hpp"

namespace qmbed {

// flip groups holding the column entries c·a of every coupling
void compile_groups(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms, const AssemblyChecks& checks,
                    bool nonempty, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off, bool& real);
// operator family: parts[0] = static terms, parts[1+p] = component p; every
// group carries one table per part (FlipGroup::parts) and amp = Σ parts
void compile_family_groups(const SpinBasisSpec& spec, const std::vector<std::vector<OperatorSpec>>& parts,
                           const AssemblyChecks& checks, bool nonempty, std::vector<FlipGroup>& diag,
                           std::vector<FlipGroup>& off, bool& real);


}  // namespace qmbed
