classdef GpuSolver
% GpuSolver: rooftopMoM.Solver (src/+rooftopMoM/Solver.m:17-92) with the Krylov loop on the GPU: the vectors, the
%   scalars and the stopping tests stay on the device and no vector crosses to the host per iteration. The operator
%   and the preconditioner may be GPU products (rooftopMoM.gpu.GpuOperator, GpuPreconditioner), or, as the reference
%   allows (Solver.m:146-154, convertToHandle.m:10-16), a matrix (copied to the device), a function handle or an object
%   with apply (called back per product). The columns of a matrix right-hand side are solved together, one Result each.
%       solver = rooftopMoM.gpu.GpuSolver('method', 'gmres', 'restart', 50, 'precision', 'c64');
%       result = solver.solve(operator, excitation, 'preconditioner', preconditioner);

    properties (Access = private)
        options_ struct;
    end

    methods
        function obj = GpuSolver(options)
        % GpuSolver: Creates the solver: the name-value inputs of rooftopMoM.Solver, and
        %       'precision': 'c128' or 'c64', the arithmetic of the products inside the loop; 'c64' runs inside an
        %           FP64 refinement loop on the true residual (default 'c128').
        %       'refine': true for that refinement loop in 'c128' too (default false).

            arguments
                options.method (1, :) char {mustBeMember(options.method, {'cg', 'gmres'})} = 'cg'
                options.tolerance (1, 1) {mustBeReal, mustBePositive, mustBeLessThan(options.tolerance, 1)} = 1e-6
                options.maxIterationCount (1, 1) {mustBeInteger, mustBePositive} = 1000
                options.restart (1, 1) {mustBePositive} = 50
                options.precision (1, :) char {mustBeMember(options.precision, {'c128', 'c64'})} = 'c128'
                options.refine (1, 1) logical = false
            end

            if ~isinf(options.restart) && options.restart ~= round(options.restart)
                error('rooftopMoM:solver:invalidRestart', 'The restart is a positive integer or Inf.');
            end

            precision = double(strcmp(options.precision, 'c64'));
            obj.options_ = struct('method', double(strcmp(options.method, 'gmres')), ...
                'tolerance', options.tolerance, 'maxIterationCount', options.maxIterationCount, ...
                'restart', options.restart, 'precision', precision, ...
                'refine', double(options.refine || precision == 1));
        end

        function result = solve(obj, operator, rightHandSide, options)
        % solve: Solves operator * x = rightHandSide on the GPU; returns a rooftopMoM.Result (an array of them for a
        %   matrix right-hand side), as rooftopMoM.Solver.solve does.

            arguments
                obj
                operator
                rightHandSide
                options.preconditioner = []
                options.initialGuess {mustBeNumeric, mustBeFinite} = []
            end

            if isa(rightHandSide, 'rooftopMoM.Excitation')
                excitation = rightHandSide;
                vectors = rightHandSide.vector();
            else
                excitation = [];
                vectors = double(rightHandSide);

                if isvector(vectors)
                    vectors = vectors(:);
                end
            end

            if ~isnumeric(vectors) || ~allfinite(vectors)
                error('rooftopMoM:solver:invalidRightHandSide', ...
                    'The right-hand side is a rooftopMoM.Excitation or a numeric vector of finite values.');
            end

            count = size(vectors, 1);
            [operatorHandle, keepOperator] = toProduct_(operator, count);
            [preconditionerHandle, keepPreconditioner] = toProduct_(options.preconditioner, count);
            initialGuess = double(options.initialGuess);

            if ~isempty(initialGuess)
                initialGuess = reshape(initialGuess, size(vectors));
            end

            [solution, outputs] = rooftop_mex('solve', rooftopMoM.gpu.context(), operatorHandle, ...
                preconditionerHandle, vectors, initialGuess, obj.options_);
            clear keepOperator keepPreconditioner

            result = rooftopMoM.Result.empty;

            for column = 1:size(vectors, 2)
                out = outputs(column);
                output = rooftopMoM.solver.Output(solution(:, column), out.converged, out.iterationCount, ...
                    out.residualNorms, out.residualNorm, out.time);
                result(column) = rooftopMoM.Result(output, excitation);
            end
        end
    end
end

function [handle, keep] = toProduct_(operator, count)
% toProduct_: The librooftop product handle of an operator (empty for none) and an object that must live while it is
%   used (a GPU product made here for a matrix or a function handle).

    keep = [];
    handle = [];

    if isempty(operator)
        return
    end

    if isa(operator, 'rooftopMoM.gpu.GpuOperator') || isa(operator, 'rooftopMoM.gpu.GpuPreconditioner')
        handle = operator.handle();
    elseif isnumeric(operator)
        if ~isequal(size(operator), [count, count])
            error('rooftopMoM:solver:sizeMismatch', 'The matrix does not match the %d unknowns.', count);
        end

        handle = rooftop_mex('matrix', rooftopMoM.gpu.context(), operator, 0);
        keep = onCleanup(@() rooftop_mex('free', 'product', handle));
    elseif isa(operator, 'function_handle') || ismethod(operator, 'apply')
        if ~isa(operator, 'function_handle')
            object = operator;
            operator = @(x) object.apply(x);
        end

        handle = rooftop_mex('handle', rooftopMoM.gpu.context(), count, operator);
        keep = onCleanup(@() rooftop_mex('free', 'product', handle));
    else
        error('rooftopMoM:solver:invalidOperator', ['An operator is a matrix, a function handle or an object with ' ...
            'the method apply.']);
    end
end
