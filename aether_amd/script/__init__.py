"""Aether's script layer over the engine: a Lua 5.4 subset interpreter
(:mod:`.lua`) and the simulation/structure/material bindings (:mod:`.front`)."""
from .front import Front, run_script  # noqa: F401
from .lua import Interpreter, LuaError  # noqa: F401
