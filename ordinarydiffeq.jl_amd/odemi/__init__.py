"""odemi -- MI355X-native ensemble ODE engine (Tsit5 / Rodas5P) behind the
reference's EnsembleProblem surface.  See api.py for the user entry points."""
from .api import (CallbackSet, ContinuousCallback, DEStats, DiscreteCallback, EnsembleContext, EnsembleMI355X, EnsembleProblem, EnsembleSolution, MixedPrecisionWarning, ODEFunction,
                  ODEProblem, ODESolution, ReturnCode, Rodas5P, Tsit5, compile_problem, get_context, remake,
                  solve, solve_arrays, terminate, VectorContinuousCallback)
from .serving import EnsembleServer, Request
from .tracer import Sym, sym_all, sym_any, sym_ifelse, sym_in, sym_cos, sym_exp, sym_log, sym_sin, sym_sqrt, sym_tanh, trace_problem

__all__ = ["CallbackSet", "ContinuousCallback", "DiscreteCallback", "VectorContinuousCallback", "sym_ifelse", "terminate", "sym_in", "DEStats", "EnsembleContext", "EnsembleMI355X", "EnsembleProblem", "EnsembleSolution", "MixedPrecisionWarning", "ODEFunction",
           "ODEProblem", "ODESolution", "ReturnCode", "Rodas5P", "Tsit5", "compile_problem", "get_context", "remake",
           "solve", "solve_arrays", "Sym", "sym_all", "sym_any", "sym_cos", "sym_exp", "sym_log", "sym_sin",
           "sym_sqrt", "sym_tanh",
           "trace_problem", "EnsembleServer", "Request"]
